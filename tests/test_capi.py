"""The C-ABI library loads and exports every symbol include/summit_gpe.h declares
(no GPU needed: no compute entry point is called with valid work)."""
from __future__ import annotations

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "summit_gpe.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpe_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = _declared()
    for must in ("gpe_phase1_xz", "gpe_phase2_accum", "gpe_gz_i32", "gpe_gtu_i64", "gpe_geno_prepare",
                 "gpe_resvar_inv", "gpe_probes", "gpe_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from summit_amd import _native
    lib = _native.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert set(_declared()) <= set(_native.SIGNATURES)


def test_error_path_reports_message_without_device_work():
    from summit_amd import _native
    lib = _native.load()
    rc = lib.gpe_geno_prepare(None, 1, 0, 0, 0, None, None)
    assert rc != 0
    assert b"null genotype handle" in lib.gpe_last_error()
    rc = lib.gpe_geno_prepare(None, 1, 7, 0, 0, None, None)
    assert rc != 0
    with pytest.raises(_native.GpeError, match="null genotype handle"):
        _native.call("gpe_resvar_inv", None, 0, 1, 1e-10, None, None)


def test_version_string():
    from summit_amd import _native
    assert b"gfx950" in _native.load().gpe_version()


def test_schedule_table_set_get_clear_and_validation():
    from summit_amd import _native, autotune
    key = autotune.shape_key(1000, 5000, 2, 40, 5)
    assert key == (1000, 5000, 128, 5)
    autotune.clear()
    assert autotune.get_value("p1_splits", key) == 0
    autotune.set_value("p1_splits", key, 4)
    autotune.set_value("p1_finalize_kib", key, 1024)
    assert autotune.get_value("p1_splits", key) == 4
    assert autotune.get_value("p1_finalize_kib", key) == 1024
    assert autotune.get_value("p1_splits", (1000, 5000, 128, 6)) == 0
    autotune.set_value("p1_splits", key, 0)  # removes the entry
    assert autotune.get_value("p1_splits", key) == 0
    with pytest.raises(_native.GpeError, match="splits"):
        autotune.set_value("p1_splits", key, 65)
    with pytest.raises(_native.GpeError, match="finalize range"):
        autotune.set_value("p1_finalize_kib", key, 8)
    with pytest.raises(_native.GpeError, match="chunks per barrier"):
        autotune.set_value("p2_cpb", key, 3)
    autotune.set_value("p1_cpb", key, 2)
    assert autotune.get_value("p1_cpb", key) == 2
    with pytest.raises(_native.GpeError, match="unknown schedule knob"):
        _native.call("gpe_schedule_set", 7, 1, 1, 1, 1, 1)
    autotune.clear()
    assert autotune.get_value("p1_finalize_kib", key) == 0
