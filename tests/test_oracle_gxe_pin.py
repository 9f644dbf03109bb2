"""Independent pins of the oracle's GxE feature construction (no GPU).

The reference's own GxE tests build their expected features with a dense
formula (tests/test_gxe_native_core.py:95-123): mean-fill the missing calls,
standardize each column with the ddof standard deviation of the filled
column, project with p = I - q q^T for an orthonormal basis q that holds the
intercept, and rescale every column so that ||x_j||^2 / rank = 1 with
rank = N - q.shape[1].  `_features_dense` below restates that formula with
NaN for missing calls; the oracle (oracle.gxe_project_center plus the
inv_x / inv_w of oracle.gxe_scores) must agree with it.  The oracle
standardizes with the observed-call variance (genotype.cpp), the reference
test with the variance of the filled column: the two differ by a per-SNP
constant, which the unit-norm rescaling removes, so the normalized features
are compared directly.
"""
from __future__ import annotations

import numpy as np

from tests.helpers import random_codes


def _features_dense(raw, env, q, ddof=1):
    filled = np.array(raw, dtype=np.float64, copy=True)
    means = np.nanmean(filled, axis=0)
    r, c = np.where(np.isnan(filled))
    filled[r, c] = means[c]
    filled -= means
    filled /= filled.std(axis=0, ddof=ddof)
    p = np.eye(filled.shape[0]) - q @ q.T
    rank = filled.shape[0] - q.shape[1]
    x = p @ filled
    w = p @ (env[:, None] * filled)
    x *= np.sqrt(rank / np.sum(x * x, axis=0))
    w *= np.sqrt(rank / np.sum(w * w, axis=0))
    return x, w, rank


def _design(n, p, seed):
    rng = np.random.default_rng(seed)
    env = rng.standard_normal(n)
    env = (env - env.mean()) / env.std(ddof=1)
    cov = rng.standard_normal((n, p))
    cov -= cov.mean(axis=0)
    C, _ = np.linalg.qr(cov)
    q, _ = np.linalg.qr(np.column_stack([np.ones(n), C]))
    return env, C, q


def _raw(codes):
    from oracle import oracle as o
    g = o.codes_to_geno(codes).astype(np.float64).T  # (N, M), -9 = missing
    return np.where(g == -9, np.nan, g)


def _oracle_features(oracle, codes, env, C):
    n = codes.shape[1]
    bed = oracle.pack_codes(codes)
    G = oracle.read_block_standardized(bed, n, 0, codes.shape[0])
    X = oracle.gxe_project_center(G, C)
    W = oracle.gxe_project_center(env[:, None] * G, C)
    return bed, X, W


def test_gxe_features_match_reference_test_formula(oracle):
    codes = random_codes(23, 70, seed=31, miss_rate=0.05)
    env, C, q = _design(70, 3, seed=9)
    bed, X, W = _oracle_features(oracle, codes, env, C)
    out = oracle.gxe_scores(bed, 70, 23, env, C, None, nvecs=2, seed=3, step_size=10)
    x_ref, w_ref, rank = _features_dense(_raw(codes), env, q)
    assert rank == 70 - 3 - 1
    # the oracle's inv_x / inv_w normalize its projected features onto the formula's
    np.testing.assert_allclose(X * out["inv_x"][None, :], x_ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(W * out["inv_w"][None, :], w_ref, rtol=0, atol=1e-12)


def test_gxe_features_invariants(oracle):
    codes = random_codes(15, 64, seed=32, miss_rate=0.08)
    env, C, _ = _design(64, 2, seed=10)
    bed, X, W = _oracle_features(oracle, codes, env, C)
    out = oracle.gxe_scores(bed, 64, 15, env, C, None, nvecs=1, seed=3, step_size=8)
    Xn = X * out["inv_x"][None, :]
    Wn = W * out["inv_w"][None, :]
    df = 64 - 2 - 1
    ones = np.ones(64)
    for F in (Xn, Wn):
        # orthogonal to the covariates and to the intercept
        assert np.max(np.abs(C.T @ F)) < 1e-12
        assert np.max(np.abs(ones @ F)) < 1e-12
        np.testing.assert_allclose(np.sum(F * F, axis=0) / df, 1.0, rtol=1e-13)
    # projection is idempotent: projecting again changes nothing
    np.testing.assert_allclose(oracle.gxe_project_center(X, C), X, rtol=0, atol=1e-12)


def test_gxe_features_catch_planted_mistakes(oracle):
    """Plausible mistakes in the feature construction no longer match the
    reference test's formula: the interaction column projected without the
    centring, the environment applied after the projection instead of before
    it, and the rank taken as N - p (the intercept's degree of freedom
    dropped)."""
    codes = random_codes(12, 50, seed=33, miss_rate=0.04)
    env, C, q = _design(50, 2, seed=11)
    bed = oracle.pack_codes(codes)
    G = oracle.read_block_standardized(bed, 50, 0, 12)
    x_ref, w_ref, rank = _features_dense(_raw(codes), env, q)

    def norm(F, r=rank):
        return F * np.sqrt(r / np.sum(F * F, axis=0))

    EG = env[:, None] * G
    assert np.max(np.abs(norm(oracle.gxe_project_center(EG, C)) - w_ref)) < 1e-12
    no_centre = EG - C @ (C.T @ EG)
    assert np.max(np.abs(norm(no_centre) - w_ref)) > 1e-3
    env_after = env[:, None] * oracle.gxe_project_center(G, C)
    assert np.max(np.abs(norm(env_after) - w_ref)) > 1e-3
    X = oracle.gxe_project_center(G, C)
    assert np.max(np.abs(norm(X, rank + 1) - x_ref)) > 1e-3


def test_gxe_scores_expectation_from_reference_formula(oracle):
    """The oracle's scores against the exact expectation built from the
    reference test's features: E[XX_jk] = sum_i (x_j.x_i / rank)^2 a_ik for
    Rademacher probes (and likewise XW, WX, WW), within 5 Monte Carlo SE."""
    codes = random_codes(10, 40, seed=8, miss_rate=0.02)
    env, C, q = _design(40, 1, seed=12)
    bed = oracle.pack_codes(codes)
    nvecs = 3000
    out = oracle.gxe_scores(bed, 40, 10, env, C, None, nvecs=nvecs, seed=5, step_size=4, vtiles=[1500, 1500])
    x, w, rank = _features_dense(_raw(codes), env, q)
    for name, L, R in (("xx", x, x), ("xw", x, w), ("wx", w, x), ("ww", w, w)):
        c = (L.T @ R) / rank
        c2 = c * c
        mean = c2.sum(axis=1)
        var = 2.0 * (mean ** 2 - (c2 * c2).sum(axis=1)) / nvecs
        assert np.all(np.abs(out[name][:, 0] - mean) <= 5.0 * np.sqrt(var) + 1e-12), name


def test_inv_resvar_is_one_on_complete_standardized_data(oracle):
    """No missing calls, ddof = 1: ||x_j||^2 = N - 1, so inv_resvar is exactly 1
    up to rounding (src/native/gwldcore.cpp:3220-3333 with df = N - 1)."""
    codes = random_codes(20, 33, seed=34, miss_rate=0.0)
    bed = oracle.pack_codes(codes)
    X = oracle.read_block_standardized(bed, 33, 0, 20, ddof=1)
    inv = oracle.inv_resvar(X)
    poly = np.std(X, axis=0) > 0
    np.testing.assert_allclose(inv[poly], 1.0, rtol=0, atol=1e-14)
