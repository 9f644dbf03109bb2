"""GRM moments tr(K^p) by sample-space probes on the GPU (summit_amd.kmoments)
against the dense GRM of the oracle's standardized genotypes, and the Mailman
GRM product of the drop-in (HWE-completed genotypes)."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu


def _write(oracle, codes, name="km"):
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, name)
    oracle.write_bed(prefix, codes)
    return prefix


def _dense_grm(oracle, codes, n, m, mode, seed, C=None, R=None):
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, mode, seed)
    if C is not None:
        Xp = X - C @ (R @ X)
        ss = np.einsum("ij,ij->j", Xp, Xp)
        inv = 1.0 / np.sqrt(np.maximum(ss / (n - C.shape[1] - 1), 1e-10))
    else:
        inv = oracle.inv_resvar(X)
    Xs = X * inv[None, :]
    K = Xs @ Xs.T / m
    if C is not None:
        P = np.eye(n) - C @ R
        K = P @ K @ P
    return K, inv


@pytest.mark.parametrize("mode,p", [("mean", 0), ("mean", 3), ("hwe", 0)])
def test_probe_products_match_dense_grm(oracle, mode, p):
    from summit_amd import gwldcore as gw, kmoments
    m, n, q = 700, 240, 12
    print(f"kmoments probe products M={m} N={n} q={q} mode={mode} p={p}", flush=True)
    codes = random_codes(m, n, seed=31 + p, miss_rate=0.02, monomorphic=1)
    prefix = _write(oracle, codes)
    C = R = None
    if p:
        Q, _ = np.linalg.qr(np.column_stack([np.ones(n), np.random.default_rng(3).standard_normal((n, p - 1))]))
        C, R = np.asfortranarray(Q), np.ascontiguousarray(Q.T)
    K, inv = _dense_grm(oracle, codes, n, m, mode, 9, C, R)
    g, _, _ = gw._geno(prefix, prefix + ".fam", None, 1, mode, 9)
    z = kmoments.sample_probes(n, q, "rademacher", 123)
    k = kmoments.probe_products(g, z, inv, m, 250, C, R)
    Y1 = K @ z
    Y2 = K @ Y1
    ref = [np.sum(z * Y1, 0), np.sum(Y1 * Y1, 0), np.sum(Y1 * Y2, 0), np.sum(Y2 * Y2, 0)]
    for got, want in zip(k, ref):
        np.testing.assert_allclose(got, want, rtol=1e-10)


def test_kmoment_traces_within_probe_error(oracle):
    from summit_amd import kmoments
    m, n, q = 900, 200, 256
    codes = random_codes(m, n, seed=5, miss_rate=0.01)
    prefix = _write(oracle, codes)
    K, inv = _dense_grm(oracle, codes, n, m, "mean", 0)
    lam = np.linalg.eigvalsh(K)
    exact = [float(np.sum(lam ** p)) for p in (1, 2, 3, 4)]
    out = str(tempfile.mkdtemp()) + "/run"
    row = kmoments.unpartitioned_kmoments(prefix, prefix + ".fam", m, 300, inv, exact[1], float(n - 1), 2024,
                                          num_probes=q, impute_mode="mean", outpath=out)
    for p, key in zip((1, 2, 3, 4), ("trace_K_probe", "trace_K2_probe", "trace_K3_probe", "trace_K4_probe")):
        se = row[key + "_se"]
        print(f"tr(K^{p}) probe {row[key]:.6g} exact {exact[p - 1]:.6g} se {se:.3g}", flush=True)
        assert abs(row[key] - exact[p - 1]) <= 5.0 * se
    assert row["probe_seed"] == kmoments.probe_seed(2024)
    assert os.path.exists(out + ".gw.kmoments")


def test_mailman_grm_product_matches_dense_hwe(oracle):
    from summit_amd import gwldcore as gw
    m, n, q = 400, 150, 3
    codes = random_codes(m, n, seed=77, miss_rate=0.03)
    prefix = _write(oracle, codes)
    K, inv = _dense_grm(oracle, codes, n, m, "hwe", 42)
    pin = np.asfortranarray(np.random.default_rng(0).standard_normal((n, q)))
    pout = np.zeros((n, q), order="F")
    gw.apply_grm_bed_panel_mailman(prefix, prefix + ".fam", m, 128, None, 1, inv_all=inv, panel_in=pin,
                                   panel_out=pout, impute_seed=42)
    np.testing.assert_allclose(pout, K @ pin, rtol=1e-10, atol=1e-12 * np.max(np.abs(K @ pin)))
    with pytest.raises(RuntimeError, match="panel row mismatch in apply_grm_bed_panel_mailman"):
        gw.apply_grm_bed_panel_mailman(prefix, prefix + ".fam", m, 128, None, 1, inv_all=inv,
                                       panel_in=pin[:-1], panel_out=pout[:-1], impute_seed=42)
