"""Pins of the oracle's windowed LD restatement (no GPU)."""
from __future__ import annotations

import numpy as np

from tests.helpers import random_codes


def test_windowed_matches_textbook_complete_data(oracle):
    # tests/test_pgen_gwld.py:899-916 without covariates: ddof-0 standardization,
    # r = x^T x / N, unbiased r^2 with corr_dim = N - 1, |dbp| <= window
    codes = random_codes(30, 45, seed=4, miss_rate=0.0)
    g = oracle.codes_to_geno(codes).astype(np.float64).T
    x = g - g.mean(axis=0)
    sd = x.std(axis=0, ddof=0)
    x = np.where(sd > 0, x / np.where(sd > 0, sd, 1.0), 0.0)
    corr = x.T @ x / 45.0
    unb = corr * corr - (1 - corr * corr) / (45 - 1 - 1)
    bp = np.cumsum(np.random.default_rng(0).integers(50, 150, size=30))
    mask = np.abs(bp[:, None] - bp[None, :]) <= 400
    annot = np.column_stack([np.ones(30), np.arange(30) % 2])
    expect = (unb * mask) @ annot
    got = oracle.windowed_ld_chr(oracle.pack_codes(codes), 45, 0, 30, bp, annot, ld_wind_kb=0.4)
    np.testing.assert_allclose(got, expect, rtol=1e-12, atol=1e-12)


def test_windowed_mean_imputation_scale(oracle):
    # missing calls: x = (g - mu_obs) o / sqrt(sum_obs (g - mu_obs)^2 / N), restandardized over all N
    codes = random_codes(12, 60, seed=5, miss_rate=0.1)
    g = oracle.codes_to_geno(codes).astype(np.float64).T
    obs = g != -9
    x = np.zeros_like(g)
    for j in range(12):
        v = g[obs[:, j], j]
        c = v - v.mean()
        x[obs[:, j], j] = c / np.sqrt((c @ c) / 60.0)
    corr = x.T @ x / 60.0
    unb = corr * corr - (1 - corr * corr) / (60 - 2)
    got = oracle.windowed_ld_chr(oracle.pack_codes(codes), 60, 0, 12, None, np.ones((12, 1)), window_snps=3)
    idx = np.arange(12)
    expect = (unb * (np.abs(idx[:, None] - idx[None, :]) <= 3)) @ np.ones((12, 1))
    np.testing.assert_allclose(got, expect, rtol=1e-11, atol=1e-12)


def test_gxe_oracle_within_five_se_of_kernel_expectation(oracle):
    """E[XX_jk] = sum_i (X_j.X_i)^2 a_ik / df^2 (and likewise XW, WX, WW) for Rademacher
    probes; the oracle's estimate must lie within 5 Monte Carlo SE."""
    codes = random_codes(10, 40, seed=8, miss_rate=0.02)
    bed = oracle.pack_codes(codes)
    rng = np.random.default_rng(3)
    env = rng.standard_normal(40)
    raw = rng.standard_normal((40, 1))
    raw -= raw.mean(axis=0)
    C, _ = np.linalg.qr(raw)
    nvecs = 3000
    out = oracle.gxe_scores(bed, 40, 10, env, C, None, nvecs=nvecs, seed=5, step_size=4, vtiles=[1500, 1500])
    G = oracle.read_block_standardized(bed, 40, 0, 10)
    df = 40 - 1 - 1
    X = oracle.gxe_project_center(G, C)
    W = oracle.gxe_project_center(env[:, None] * G, C)
    X /= np.sqrt((X * X).sum(0) / df)
    W /= np.sqrt((W * W).sum(0) / df)
    # the normalized features have norm^2 / df == 1 (gwe_ldscore.py:2440-2458)
    np.testing.assert_allclose((X * X).sum(0) / df, 1.0, rtol=1e-12)
    for name, L, R in (("xx", X, X), ("xw", X, W), ("wx", W, X), ("ww", W, W)):
        c = (L.T @ R) / df
        c2 = c * c
        mean = c2.sum(axis=1)
        var = 2.0 * (mean ** 2 - (c2 * c2).sum(axis=1)) / nvecs
        assert np.all(np.abs(out[name][:, 0] - mean) <= 5.0 * np.sqrt(var) + 1e-12), name


def test_windowed_covariates_match_reference_test_formula(oracle):
    """tests/test_pgen_gwld.py:899-916: project the ddof-0 standardized genotypes
    off q = QR([1, covariate]), re-centre and rescale to N, corr_dim = N - q.shape[1]."""
    rng = np.random.default_rng(7)
    codes = random_codes(25, 40, seed=11, miss_rate=0.0)
    g = oracle.codes_to_geno(codes).astype(np.float64).T
    raw = g - g.mean(axis=0)
    raw /= raw.std(axis=0, ddof=0)
    age = rng.standard_normal(40)
    q, _ = np.linalg.qr(np.column_stack([np.ones(40), age]))
    x = raw - q @ (q.T @ raw)
    x -= x.mean(axis=0)
    x *= np.sqrt(40.0 / np.sum(x * x, axis=0))
    corr = x.T @ x / 40.0
    unb = corr * corr - (1.0 - corr * corr) / float(40 - 2 - 1)
    bp = np.cumsum(rng.integers(20, 80, size=25))
    annot = np.column_stack([np.ones(25), (np.arange(25) % 3 == 0).astype(float)])
    expect = (unb * (np.abs(bp[:, None] - bp[None, :]) <= 100)) @ annot
    got = oracle.windowed_ld_chr(oracle.pack_codes(codes), 40, 0, 25, bp, annot, ld_wind_kb=0.1, C=q, R=q.T)
    np.testing.assert_allclose(got, expect, rtol=2e-12, atol=2e-12)
    # a dropped projection or the uncorrected degrees of freedom no longer match
    plain = oracle.windowed_ld_chr(oracle.pack_codes(codes), 40, 0, 25, bp, annot, ld_wind_kb=0.1)
    assert np.max(np.abs(plain - expect)) > 1e-3


def test_windowed_covariates_missing_calls_general_factor(oracle):
    """Missing calls (mean imputation) and a non-orthonormal C with R = (C^T C)^-1 C^T:
    the oracle equals the textbook residual correlation of mean-imputed genotypes."""
    rng = np.random.default_rng(8)
    codes = random_codes(18, 60, seed=12, miss_rate=0.1)
    g = oracle.codes_to_geno(codes).astype(np.float64).T
    obs = g != -9
    x = np.zeros_like(g)
    for j in range(18):
        v = g[obs[:, j], j]
        c = v - v.mean()
        x[obs[:, j], j] = c / np.sqrt((c @ c) / obs[:, j].sum())
    C = np.column_stack([np.ones(60), rng.standard_normal((60, 2)) * 3.0 + 1.0])
    R = np.linalg.solve(C.T @ C, C.T)
    y = x - C @ (R @ x)
    y -= y.mean(axis=0)
    y *= np.sqrt(60.0 / np.sum(y * y, axis=0))
    corr = y.T @ y / 60.0
    unb = corr * corr - (1.0 - corr * corr) / float(60 - 3 - 1)
    expect = (unb * (np.abs(np.arange(18)[:, None] - np.arange(18)[None, :]) <= 4)) @ np.ones((18, 1))
    got = oracle.windowed_ld_chr(oracle.pack_codes(codes), 60, 0, 18, None, np.ones((18, 1)), window_snps=4, C=C,
                                 R=R)
    np.testing.assert_allclose(got, expect, rtol=1e-11, atol=1e-12)
