"""Dense standardized products X.C, X^T V and the GRM panel apply against the oracle."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("mode,limbs,tol", [("mean", 5, 1e-10), ("hwe", 6, 1e-12)])
def test_xc_and_xtv_match_dense(oracle, mode, limbs, tol):
    from summit_amd import engine
    m, n, q = 600, 347, 9
    codes = random_codes(m, n, seed=17, miss_rate=0.03, monomorphic=1)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(impute=mode, impute_seed=3)
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, mode, 3)
    rng = np.random.default_rng(0)
    c = rng.standard_normal((m, q))
    lo, hi = 50, 555
    out = engine.xc(g, torch.from_numpy(c[lo:hi]).cuda(), lo, hi, limbs=limbs).cpu().numpy().T
    ref = X[:, lo:hi] @ c[lo:hi]
    assert np.max(np.abs(out - ref)) <= tol * np.max(np.abs(ref))
    v = rng.standard_normal((n, q))
    w = engine.xtv(g, torch.from_numpy(np.ascontiguousarray(v.T)).cuda(), lo, hi, limbs=limbs).cpu().numpy()
    ref2 = X[:, lo:hi].T @ v
    assert np.max(np.abs(w - ref2)) <= tol * np.max(np.abs(ref2))
    g.close()


def test_apply_grm_panel_matches_dense(oracle):
    from summit_amd import gwldcore as gw
    m, n, q = 500, 260, 4
    codes = random_codes(m, n, seed=23, miss_rate=0.02)
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, "grm")
    oracle.write_bed(prefix, codes)
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, "mean", 0)
    inv = oracle.inv_resvar(X)
    rng = np.random.default_rng(1)
    pin = np.asfortranarray(rng.choice([-1.0, 1.0], size=(n, q)))
    pout = np.zeros((n, q), order="F")
    gw.apply_grm_bed_panel(prefix, prefix + ".fam", 450, 128, None, 1, inv_all=inv, panel_in=pin, panel_out=pout,
                           impute_mode="mean")
    Xs = X[:, :450] * inv[None, :450]
    ref = Xs @ (Xs.T @ pin) / 450.0
    assert np.max(np.abs(pout - ref)) <= 1e-11 * np.max(np.abs(ref))


@pytest.mark.parametrize("p,center", [(0, True), (3, False), (3, True), (40, True)])
def test_projection_kernel_matches_numpy(p, center):
    """x <- x - A (B x) [- mean] in FP64 on the GPU (gpe_project_f64), including
    more factor rows than one coefficient pass (40 > 32) and centring only."""
    from summit_amd import engine
    rng = np.random.default_rng(p + 10 * center)
    Q, N = 7, 1000
    x = rng.standard_normal((Q, N))
    A = rng.standard_normal((N, p))
    B = rng.standard_normal((p, N)) / N
    ref = x.T - A @ (B @ x.T) if p else x.T.copy()
    if center:
        ref = ref - ref.mean(axis=0, keepdims=True)
    xd = torch.from_numpy(x).cuda()
    if p:
        engine.project(xd, torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda(), center=center)
    else:
        engine.project(xd, None, None, center=center)
    np.testing.assert_allclose(xd.cpu().numpy(), ref.T, rtol=1e-12, atol=1e-12)
