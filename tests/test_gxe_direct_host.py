"""Host-side checks of the gxeldcore.DirectContext drop-in that run before any
GPU work: argument, descriptor, file-format, row-selection and design
validation with the reference's error texts (src/native/gxeldcore.cpp:7086-7200,
:6460-6553, :8504-8655), and the opaque ProjectedPanel."""
from __future__ import annotations

import os

import numpy as np
import pytest

from summit_amd import gxeldcore


def _write(oracle, tmp_path, n=12, m=5, name="d"):
    rng = np.random.default_rng(3)
    g = rng.integers(0, 3, size=(m, n))
    prefix = str(tmp_path / name)
    oracle.write_bed(prefix, oracle.geno_to_codes(g))
    return prefix


def _fds(prefix):
    return [os.open(prefix + ext, os.O_RDONLY | os.O_CLOEXEC) for ext in (".bed", ".bim", ".fam")]


def _design(n, ddof=1):
    rng = np.random.default_rng(11)
    env = rng.normal(size=n)
    env = (env - env.mean()) / env.std(ddof=ddof)
    q, _ = np.linalg.qr(np.column_stack([np.ones(n), env, rng.normal(size=n)]))
    return env, np.asfortranarray(q)


def _ctx(fds, env, q, **kw):
    args = dict(row_sel=None, ddof=1, env=env, q_basis=q, decode_threads=1, max_workspace_bytes=1 << 30)
    args.update(kw)
    return gxeldcore.DirectContext(fds[0], fds[1], fds[2], **args)


@pytest.fixture
def files(oracle, tmp_path):
    prefix = _write(oracle, tmp_path)
    fds = _fds(prefix)
    yield prefix, fds
    for fd in fds:
        os.close(fd)


@pytest.mark.parametrize("kw,msg", [
    ({"ddof": 2}, "supports ddof 0 or 1"),
    ({"decode_threads": 0}, "decode_threads must be positive"),
    ({"decode_threads": 10 ** 6}, "exceeds the OpenMP/CPU-affinity limit"),
    ({"max_workspace_bytes": 0}, "max_workspace_bytes is invalid"),
    ({"target_panel_columns": 0}, "target_panel_columns must be positive"),
])
def test_argument_checks(files, kw, msg):
    _, fds = files
    env, q = _design(12)
    with pytest.raises(RuntimeError, match=msg):
        _ctx(fds, env, q, **kw)


def test_descriptor_and_format_checks(oracle, tmp_path, files):
    prefix, fds = files
    env, q = _design(12)
    with pytest.raises(RuntimeError, match="Invalid GxE native BED descriptor"):
        gxeldcore.DirectContext(-1, fds[1], fds[2], None, 1, env, q, 1, 1 << 30)
    # a BED whose size disagrees with FAM x BIM
    short = str(tmp_path / "short")
    for ext in (".bim", ".fam"):
        with open(prefix + ext, "rb") as f, open(short + ext, "wb") as o:
            o.write(f.read())
    with open(prefix + ".bed", "rb") as f, open(short + ".bed", "wb") as o:
        o.write(f.read()[:-1])
    sfds = _fds(short)
    try:
        with pytest.raises(RuntimeError, match="BED byte size disagrees with FAM/BIM dimensions"):
            _ctx(sfds, env, q)
    finally:
        for fd in sfds:
            os.close(fd)
    # bad magic
    with open(prefix + ".bed", "rb") as f:
        body = bytearray(f.read())
    body[2] = 0
    with open(short + ".bed", "wb") as o:
        o.write(bytes(body))
    sfds = _fds(short)
    try:
        with pytest.raises(RuntimeError, match="not a SNP-major PLINK BED"):
            _ctx(sfds, env, q)
    finally:
        for fd in sfds:
            os.close(fd)


def test_row_counting_rules(tmp_path):
    def count(text: bytes):
        p = tmp_path / "rows.txt"
        p.write_bytes(text)
        fd = os.open(str(p), os.O_RDONLY)
        try:
            return gxeldcore._count_rows(fd, len(text), "FAM")
        finally:
            os.close(fd)

    assert count(b"a b\nc d\n") == 2
    assert count(b"a b\nc d") == 2  # a final row without a newline counts
    assert count(b"a\r\n\tb\r\n") == 2
    with pytest.raises(RuntimeError, match="contains a blank row"):
        count(b"a\n \t\nb\n")
    with pytest.raises(RuntimeError, match="contains a blank final row"):
        count(b"a\nb\n  ")
    with pytest.raises(RuntimeError, match="has invalid row count"):
        count(b"")


@pytest.mark.parametrize("rows,msg", [
    (np.array([0, 1, 2 ** 32], dtype=np.int64), "out-of-range sample index"),
    (np.array([0, 2, 2, 5], dtype=np.int64), "strictly increasing and within the FAM axis"),
    (np.array([0, 5, 12], dtype=np.int32), "strictly increasing and within the FAM axis"),
    (np.array([0, 1], dtype=np.int64), "at least three selected samples"),
    (np.array([0.0, 1.0, 2.0]), "dtype int32 or int64"),
    (np.zeros((2, 2), dtype=np.int64), "bounded one-dimensional integer array"),
])
def test_row_selection_checks(files, rows, msg):
    _, fds = files
    n = rows.shape[0] if rows.ndim == 1 else 4
    env, q = _design(max(n, 3))
    with pytest.raises(RuntimeError, match=msg):
        _ctx(fds, env[:n] if n >= 3 else env, q[:n] if n >= 3 else q, row_sel=rows)


def test_design_checks(files):
    _, fds = files
    env, q = _design(12)
    with pytest.raises(RuntimeError, match="environment/projection row mismatch"):
        _ctx(fds, env[:11], q)
    with pytest.raises(RuntimeError, match="projection basis has invalid rank"):
        _ctx(fds, env, np.zeros((12, 0)))
    bad = env.copy()
    bad[3] = np.inf
    with pytest.raises(RuntimeError, match="environment contains a non-finite value"):
        _ctx(fds, bad, q)
    with pytest.raises(RuntimeError, match="centered, and standardized for the configured ddof"):
        _ctx(fds, env * 1.5, q)
    # the ddof the environment was standardized for matters
    with pytest.raises(RuntimeError, match="centered, and standardized for the configured ddof"):
        _ctx(fds, env, q, ddof=0)
    with pytest.raises(RuntimeError, match="projection basis is not orthonormal"):
        _ctx(fds, env, q * 1.01)
    # orthonormal, but missing the environment direction
    rng = np.random.default_rng(2)
    q2, _ = np.linalg.qr(np.column_stack([np.ones(12), rng.normal(size=(12, 2))]))
    with pytest.raises(RuntimeError, match="must span the intercept and environment"):
        _ctx(fds, env, q2)
    qn = q.copy()
    qn[0, 0] = np.nan
    with pytest.raises(RuntimeError, match="projection basis contains a non-finite value"):
        _ctx(fds, env, qn)


def test_projected_panel_has_no_public_constructor():
    with pytest.raises(TypeError):
        gxeldcore.ProjectedPanel()
