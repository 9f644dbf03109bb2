"""The cross-trait moment solve (summit_amd.rg) against closed forms, the
univariate solve it must reduce to, and simulated GWAS pairs whose genetic
correlation is known (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from summit_amd import h2, rg


def test_one_bin_closed_form():
    """K = 1, a_j = 1: T gamma = sqrt(n1 n2)(ybar - c) with T = n1 n2 Lbar / M gives
    gamma = M (ybar - c) / (sqrt(n1 n2) Lbar); a replicate is the same formula
    on the remaining SNPs, still scaled to the full SNP count."""
    rng = np.random.default_rng(11)
    M, n1, n2, c = 2500, 18000.0, 31000.0, 0.07
    L = rng.uniform(5.0, 60.0, size=M)
    y = 0.05 + rng.normal(0.2, 1.0, size=M)
    d = h2.JackknifeDesign.blocks(M, 25)
    y1 = 1.0 + rng.gamma(2.0, 0.5, size=M)
    f1 = h2.fit_h2(L, np.ones(M), y1, n1, d)
    f2 = h2.fit_h2(L, np.ones(M), y1 * 1.3, n2, d)
    fit = rg.fit_rg(L, np.ones(M), y, n1, n2, d, f1, f2, c=c)
    g = M * (y.mean() - c) / (np.sqrt(n1 * n2) * L.mean())
    assert fit.gamma[0] == pytest.approx(g, rel=1e-12)
    assert fit.gamma_total == pytest.approx(g, rel=1e-12)
    keep = np.ones(M, bool)
    keep[d.starts[7]:d.ends[7]] = False
    g7 = M * (y[keep].mean() - c) / (np.sqrt(n1 * n2) * L[keep].mean())
    assert fit.gamma_reps[7, 0] == pytest.approx(g7, rel=1e-11)
    # per replicate rg scales by that replicate's univariate components
    assert fit.rg_reps[7, 0] == pytest.approx(g7 / np.sqrt(f1.sigma_reps[7, 0] * f2.sigma_reps[7, 0]), rel=1e-11)
    assert fit.rg_total == pytest.approx(g / np.sqrt(f1.h2 * f2.h2), rel=1e-12)


def test_a_trait_with_itself_is_the_univariate_solve():
    """y1 = y2 with complete sample sizes and c = 1: eliminating the noise
    component from the univariate system [T + n, n; n, n] leaves exactly
    T gamma = n (ybar - 1), so every replicate's gamma equals the univariate
    sigma and every rg is one (several overlapping bins, symmetrized traces)."""
    rng = np.random.default_rng(12)
    M, K, n = 3000, 3, 25000.0
    A = np.column_stack([np.ones(M), (rng.random(M) < 0.3), rng.uniform(0.0, 2.0, M)]).astype(float)
    L = A * rng.uniform(3.0, 40.0, size=(M, 1)) + rng.normal(0, 2.0, size=(M, K))
    beta = rng.normal(0, 0.01, M)
    se = np.full(M, 1.0 / np.sqrt(n))
    y1 = h2.h2_response(beta, se, n + 1, n + 1)
    y, ov, n1s, n2s = rg.rg_response(beta, se, n + 1, n + 1, beta, se, n + 1, n + 1)
    np.testing.assert_allclose(y, y1, rtol=1e-13)
    np.testing.assert_allclose(ov, 1.0, rtol=1e-14)
    d = h2.JackknifeDesign.blocks(M, 30)
    f = h2.fit_h2(L, A, y1, n1s, d)
    fit = rg.fit_rg(L, A, y, n1s, n2s, d, f, f, c=1.0, overlap=ov)
    np.testing.assert_allclose(fit.gamma_reps, f.sigma_reps[:, :K], rtol=1e-8, atol=1e-12)
    pos = f.sigma_reps[:, :K] > 0
    np.testing.assert_allclose(fit.rg_reps[pos], 1.0, rtol=1e-8)
    assert np.isnan(fit.rg_reps[~pos]).all()
    assert fit.rg_total == pytest.approx(1.0, rel=1e-9)
    assert fit.gamma_total == pytest.approx(f.h2, rel=1e-9)
    assert fit.gamma_total_se == pytest.approx(f.h2_se, rel=1e-6)


def test_overlap_multiplier_and_c_are_linear():
    """The supplied covariance enters through c * C_k: a constant multiplier kappa
    equals c kappa, and gamma is affine in c."""
    rng = np.random.default_rng(13)
    M, n = 1800, 9000.0
    A = np.column_stack([np.ones(M), rng.random(M) < 0.5]).astype(float)
    L = A * rng.uniform(3.0, 30.0, size=(M, 1))
    y = rng.normal(0.3, 1.0, M)
    d = h2.JackknifeDesign.blocks(M, 20)
    f = h2.fit_h2(L, A, 1.0 + np.abs(y), n, d)
    a = rg.fit_rg(L, A, y, n, n, d, f, f, c=0.2, overlap=np.full(M, 1.5))
    b = rg.fit_rg(L, A, y, n, n, d, f, f, c=0.3)
    np.testing.assert_allclose(a.gamma_reps, b.gamma_reps, rtol=1e-12, atol=1e-15)
    g0 = rg.fit_rg(L, A, y, n, n, d, f, f, c=0.0).gamma_reps
    g1 = rg.fit_rg(L, A, y, n, n, d, f, f, c=0.6).gamma_reps
    np.testing.assert_allclose(b.gamma_reps, 0.5 * (g0 + g1), rtol=1e-12, atol=1e-15)


def _genotypes(rng, N, M):
    hap = np.empty((2, N, M), dtype=np.int8)
    maf = rng.uniform(0.05, 0.5, size=M)
    for h in range(2):
        cur = rng.random(N) < maf[0]
        for j in range(M):
            if j % 20 == 0:
                cur = rng.random(N) < maf[j]
            else:
                redraw = rng.random(N) > 0.8
                cur = np.where(redraw, rng.random(N) < maf[j], cur)
            hap[h, :, j] = cur
    return (hap[0] + hap[1]).astype(np.float64)


def _gwas(G, ph):
    X = (G - G.mean(0)) / G.std(0)
    ph = ph - ph.mean()
    N = X.shape[0]
    xx = np.sum(X * X, axis=0)
    b = X.T @ ph / xx
    resid = np.sum(ph * ph) - b * b * xx
    return b, np.sqrt(resid / (N - 2) / xx)


def _simulate_pair(rng, N, M, h2_1, h2_2, rho_g, overlap, rho_e=0.0):
    """Two traits with per-SNP effects of correlation rho_g.  overlap=False: the
    GWAS use disjoint halves of 2N samples; True: one sample of N with
    environmental correlation rho_e.  LD scores from every genotyped sample,
    null-subtracted."""
    Ntot = N if overlap else 2 * N
    G = _genotypes(rng, Ntot, M)
    X = (G - G.mean(0)) / G.std(0)
    u1, u2 = rng.standard_normal(M), rng.standard_normal(M)
    b1 = np.sqrt(h2_1 / M) * u1
    b2 = np.sqrt(h2_2 / M) * (rho_g * u1 + np.sqrt(1 - rho_g * rho_g) * u2)
    e1 = rng.standard_normal(Ntot)
    e2 = rho_e * e1 + np.sqrt(1 - rho_e * rho_e) * rng.standard_normal(Ntot)
    p1 = X @ b1 + np.sqrt(1 - h2_1) * e1
    p2 = X @ b2 + np.sqrt(1 - h2_2) * e2
    s1 = slice(0, N)
    s2 = slice(0, N) if overlap else slice(N, 2 * N)
    r = X.T @ X / Ntot
    L = (r * r).sum(axis=1, keepdims=True) - M / (Ntot - 1)
    bb1, se1 = _gwas(G[s1], p1[s1])
    bb2, se2 = _gwas(G[s2], p2[s2])
    c = float(np.corrcoef(p1[s1], p2[s1])[0, 1]) if overlap else 0.0
    return bb1, se1, bb2, se2, L, c


@pytest.mark.parametrize("rho_g,overlap", [(0.6, False), (-0.5, True)])
def test_simulated_rg_recovered_within_jackknife_se(rho_g, overlap):
    rng = np.random.default_rng(21 if overlap else 20)
    N, M = 3000, 1600
    b1, se1, b2, se2, L, c = _simulate_pair(rng, N, M, 0.5, 0.4, rho_g, overlap, rho_e=0.9)
    A = np.ones((M, 1))
    d = h2.JackknifeDesign.blocks(M, 40)
    y1 = h2.h2_response(b1, se1, N, N)
    y2 = h2.h2_response(b2, se2, N, N)
    n = h2.effective_n_scale(N)
    f1 = h2.fit_h2(L, A, y1, n, d)
    f2 = h2.fit_h2(L, A, y2, n, d)
    y, ov, n1, n2 = rg.rg_response(b1, se1, N, N, b2, se2, N, N)
    fit = rg.fit_rg(L, A, y, n1, n2, d, f1, f2, c=c, overlap=ov)
    assert 0.02 < fit.rg_total_se < 0.4
    assert abs(fit.rg_total - rho_g) < 4.0 * fit.rg_total_se
    assert abs(fit.gamma_total - rho_g * np.sqrt(0.5 * 0.4)) < 4.0 * fit.gamma_total_se
    assert np.sign(fit.rg_total) == np.sign(rho_g)
    dl = rg.fit_rg(L, A, y, n1, n2, d, f1, f2, c=c, overlap=ov, se_method="delta")
    assert dl.rg_total == fit.rg_total and dl.rg[0] == pytest.approx(fit.rg_total)
    assert 0.5 < dl.rg_total_se / fit.rg_total_se < 2.0
    if overlap:
        # ignoring the shared samples' phenotypic covariance shifts gamma by
        # c * M / (sqrt(n1 n2) Lbar) (the one-bin closed form)
        nc = rg.fit_rg(L, A, y, n1, n2, d, f1, f2, c=0.0)
        assert nc.gamma_total - fit.gamma_total == pytest.approx(
            c * M / (np.sqrt(n1 * n2) * L[np.isfinite(y), 0].mean()), rel=1e-9)


def test_input_checks():
    M = 200
    L = np.ones((M, 1))
    d = h2.JackknifeDesign.blocks(M, 10)
    f = h2.fit_h2(L, np.ones(M), np.full(M, 1.5), 1000.0, d)
    with pytest.raises(ValueError):
        rg.fit_rg(L, np.ones(M), np.ones(M), 1000.0, 1000.0, d, f, f, se_method="robust")
    with pytest.raises(ValueError):
        rg.fit_rg(L, np.ones(M), np.ones(M), 1000.0, 1000.0, h2.JackknifeDesign.blocks(M, 5), f, f)
    with pytest.raises(ValueError):
        rg.fit_rg(L[:-1], np.ones(M), np.ones(M), 1000.0, 1000.0, d, f, f)
    with pytest.raises(RuntimeError):
        rg.fit_rg(L, np.ones(M), np.ones(M), -1.0, 1000.0, d, f, f)
    with pytest.raises(ValueError):
        rg.fit_rg(L, np.ones(M), np.full(M, np.nan), 1000.0, 1000.0, d, f, f)
