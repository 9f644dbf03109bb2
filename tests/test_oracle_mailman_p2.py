"""Pins of the oracle's Mailman phase 2 with caller-supplied column sums
(oracle.phase2_accum_mailman, src/native/gwldcore.cpp:774-908, :2919-3119), no GPU."""
from __future__ import annotations

import numpy as np

from tests.helpers import random_codes


def _hwe_block(oracle, codes, n):
    bed = oracle.pack_codes(codes)
    c, nobs, s, ss = oracle.decode_block(bed, n, 0, codes.shape[0])
    X, mu, inv_std = oracle.standardize_block(c, nobs, s, ss, 1, "hwe", impute_seed=11, blk_start=0)
    comp, _, _ = oracle.hwe_impute(c, nobs, s, 11, 0)
    G = np.array([0, 0, 1, 2], dtype=np.int64)[comp].T  # (N, L) completed counts
    return X, G, mu, inv_std


def test_mailman_with_exact_sums_equals_dense_phase2(oracle):
    codes = random_codes(40, 90, seed=21, miss_rate=0.05, monomorphic=1)
    X, G, mu, inv_std = _hwe_block(oracle, codes, 90)
    # the completed counts reproduce the standardized block
    np.testing.assert_allclose((G - mu[None, :]) * inv_std[None, :], X, rtol=0, atol=1e-13)
    rng = np.random.default_rng(4)
    Y = rng.standard_normal((90, 6))
    inv = rng.uniform(0.8, 1.2, size=40)
    dense = np.zeros((40, 2))
    mc_d = np.zeros((40, 2))
    oracle.phase2_accum(np.asfortranarray(X), inv, 3, np.asfortranarray(Y), dense, 90, 0, mc_per_snp_sumsq=mc_d)
    mm = np.zeros((40, 2))
    mc_m = np.zeros((40, 2))
    oracle.phase2_accum_mailman(G, mu, inv_std, inv, 3, Y, Y.sum(axis=0), mm, 90, 0, mc_per_snp_sumsq=mc_m)
    np.testing.assert_allclose(mm, dense, rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(mc_m, mc_d, rtol=1e-10, atol=1e-18)


def test_mailman_uses_the_callers_sums_brute_force(oracle):
    codes = random_codes(7, 25, seed=22, miss_rate=0.08)
    _, G, mu, inv_std = _hwe_block(oracle, codes, 25)
    rng = np.random.default_rng(5)
    Y = rng.standard_normal((25, 4))
    s = Y.sum(axis=0) + np.array([0.5, -1.0, 2.0, 0.25])
    inv = rng.uniform(0.5, 1.5, size=7)
    got = np.zeros((7, 2))
    oracle.phase2_accum_mailman(G, mu, inv_std, inv, 2, Y, s, got, 25, 0)
    exp = np.zeros((7, 2))
    for j in range(7):
        for k in range(2):
            for v in range(2):
                c = 2 * k + v
                w = sum(G[n, j] * Y[n, c] for n in range(25)) - mu[j] * s[c]
                exp[j, k] += (w * inv_std[j] * inv[j]) ** 2 / (25 - 1) ** 2
    np.testing.assert_allclose(got, exp, rtol=1e-12)
    # the shifted sums change the answer: a caller's sums are not silently replaced
    ref = np.zeros((7, 2))
    oracle.phase2_accum_mailman(G, mu, inv_std, inv, 2, Y, Y.sum(axis=0), ref, 25, 0)
    assert np.max(np.abs(ref - got)) > 1e-3


def test_mailman_guards_skip_unscaled_snps(oracle):
    codes = random_codes(3, 20, seed=23, miss_rate=0.0)
    _, G, mu, inv_std = _hwe_block(oracle, codes, 20)
    Y = np.ones((20, 2))
    inv = np.array([1.0, 0.0, 1e12])
    out = np.zeros((3, 1))
    counts = oracle.phase2_accum_mailman(G, mu, inv_std, inv, 2, Y, np.array([np.nan, 1.0]), out, 20, 0)
    # SNP 1 (inv 0) is skipped before the guards; SNPs 0 and 2 see a NaN sum
    assert counts == {"nonfinite": 2, "extreme": 0} and np.all(out == 0.0)
    out2 = np.zeros((3, 1))
    counts2 = oracle.phase2_accum_mailman(G, mu, inv_std, inv, 2, Y * 1e3, Y.sum(axis=0) * 1e3 + 1e3, out2, 20, 0)
    assert counts2["extreme"] == 1 and out2[2, 0] == 0.0 and out2[0, 0] > 0.0
