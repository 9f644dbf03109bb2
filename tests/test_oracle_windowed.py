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
