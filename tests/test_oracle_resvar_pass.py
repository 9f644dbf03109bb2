"""The oracle's fused residual-variance pass (residual scales, mu22 and the
SNP-variance share V) pinned by hand-computed known answers, the reference
test's own formula for V (tests/test_snp_var_weights.py:52-67) and planted
mistakes; the drop-in's block-level corrections by their closed-form cases."""
from __future__ import annotations

import numpy as np
import pytest


def _bed(oracle, G):
    """G (N, M) values {0, 1, 2, nan} -> packed SNP-major codes."""
    vals = np.where(np.isnan(G), -9, G).astype(np.int64).T
    return oracle.pack_codes(oracle.geno_to_codes(vals))


def test_mu22_known_answers(oracle):
    # one SNP 0, 0, 2, 2: x^2 = 3/4 each, inv = 1, mu22 = mean x^4 = 9/16
    G = np.array([[0.0], [0.0], [2.0], [2.0]])
    _, mu, _ = oracle.resvar_pass(_bed(oracle, G), 4, 1, 1, annot=np.ones((1, 1)))
    assert mu[0, 0] == pytest.approx(9.0 / 16.0, rel=1e-14)
    # duplicating the SNP changes nothing (S doubles, Msum doubles)
    G2 = np.column_stack([G, G])
    _, mu2, _ = oracle.resvar_pass(_bed(oracle, G2), 4, 2, 1, annot=np.ones((2, 1)))
    assert mu2[0, 0] == pytest.approx(9.0 / 16.0, rel=1e-14)
    # 0, missing, 2, 2 (mean imputation): x^2 = 4/3, 0, 1/3, 1/3; ||x||^2 = 2 over N - 1 = 3
    # gives inv^2 = 3/2, S = 2, 0, 1/2, 1/2 and mu22 = (4 + 1/4 + 1/4) / 4 = 9/8
    Gm = np.array([[0.0], [np.nan], [2.0], [2.0]])
    inv, mum, v = oracle.resvar_pass(_bed(oracle, Gm), 4, 1, 1, annot=np.ones((1, 1)))
    assert inv[0] ** 2 == pytest.approx(1.5, rel=1e-14)
    assert mum[0, 0] == pytest.approx(9.0 / 8.0, rel=1e-14)
    assert v[0] == 1.0
    # planted mistakes the known answer rejects: no inv^2 weight (0.5), the sample
    # variance of observed calls only as the scale (inv^2 = 1: 0.5), ddof 0
    assert abs(mum[0, 0] - 0.5) > 0.1
    _, mu0, _ = oracle.resvar_pass(_bed(oracle, Gm), 4, 1, 1, ddof=0, annot=np.ones((1, 1)))
    assert mu0[0, 0] == pytest.approx(9.0 / 8.0, rel=1e-14)  # ddof cancels against the residual scale


def test_mu22_invariants_and_bins(oracle):
    rng = np.random.default_rng(3)
    N, M = 60, 40
    G = rng.binomial(2, rng.uniform(0.1, 0.5, M)[None, :], size=(N, M)).astype(float)
    bed = _bed(oracle, G)
    A = np.column_stack([np.ones(M), (rng.random(M) < 0.4).astype(float), rng.uniform(0, 2, M)])
    inv, mu, _ = oracle.resvar_pass(bed, N, M, 7, annot=A)
    # complete data, all polymorphic: sum_i S_i = (N - 1) M for the all-ones bin, so by
    # Cauchy-Schwarz mu22 >= ((N - 1) / N)^2
    assert mu[0, 0] >= ((N - 1) / N) ** 2
    assert np.allclose(mu, mu.T, rtol=1e-13)
    # block size does not matter
    _, mu_b, _ = oracle.resvar_pass(bed, N, M, 40, annot=A)
    np.testing.assert_allclose(mu_b, mu, rtol=1e-12)
    # an intercept-only projection removes nothing from centred genotypes but costs a
    # degree of freedom: inv^2 scales by (N - 2) / (N - 1), mu22 by its square
    c = np.full((N, 1), 1.0 / np.sqrt(N))
    inv_c, mu_c, v_c = oracle.resvar_pass(bed, N, M, 7, annot=A, C=c, R=c.T)
    np.testing.assert_allclose(inv_c ** 2, inv ** 2 * (N - 2) / (N - 1), rtol=1e-12)
    np.testing.assert_allclose(mu_c, mu * ((N - 2) / (N - 1)) ** 2, rtol=1e-12)
    np.testing.assert_allclose(v_c, 1.0, rtol=1e-12)


def test_snp_var_matches_reference_test_formula(oracle):
    """V before the adjusted-R2 factor: ||P x||^2 / ||x||^2 of mean-imputed centred
    genotypes, P the projection off [1, covariates] (tests/test_snp_var_weights.py:52-67),
    0 for a monomorphic SNP."""
    rng = np.random.default_rng(1)
    N, M = 180, 50
    pop = (rng.random(N) < 0.5).astype(float)
    f0 = rng.uniform(0.08, 0.5, M)
    f1 = np.clip(f0 + rng.normal(0, 0.15, M), 0.03, 0.97)
    G = rng.binomial(2, np.where(pop[:, None] == 1, f1[None, :], f0[None, :])).astype(np.float64)
    G[:, 7] = 0.0
    G[rng.random(G.shape) < 0.01] = np.nan
    covs = np.column_stack([pop + rng.normal(0, 0.3, N), rng.normal(size=N)])
    X = np.where(np.isnan(G), np.nanmean(G, 0)[None, :], G) - np.nanmean(G, 0)[None, :]
    Q, _ = np.linalg.qr(np.column_stack([np.ones(N), covs]))
    PX = X - Q @ (Q.T @ X)
    tot, res = (X ** 2).sum(0), (PX ** 2).sum(0)
    ok = tot > 1e-12
    expect = np.zeros(M)
    expect[ok] = res[ok] / tot[ok]
    inv, _, v = oracle.resvar_pass(_bed(oracle, G), N, M, 13, C=Q, R=Q.T)
    np.testing.assert_allclose(v, expect, rtol=1e-12, atol=1e-15)
    assert v[7] == 0.0
    # a planted mistake: the share against the projected instead of the raw norm is 1
    assert np.max(np.abs(v[ok] - 1.0)) > 1e-3
    # the residual scales over N - p - 1 with p = 3 columns of C
    Xs = oracle.read_block_standardized(_bed(oracle, G), N, 0, M)
    Xp = Xs - Q @ (Q.T @ Xs)
    ss = (Xp ** 2).sum(0)
    np.testing.assert_allclose(inv[ok], 1.0 / np.sqrt(np.maximum(ss[ok] / (N - 4), 1e-10)), rtol=1e-12)


def test_block_corrections_closed_forms():
    from summit_amd.gwldcore import compute_block_corrections_binary
    rng = np.random.default_rng(5)
    M, N, d = 30, 500, 499.0
    A = np.column_stack([np.ones(M), (rng.random(M) < 0.5).astype(float), rng.uniform(0, 3, M)])
    A[:4, 2] = 0.0
    meansq = rng.uniform(1, 5, size=(M, 3))
    mu = rng.uniform(1, 2, size=(3, 3))
    R2, rho2, bias, delta = compute_block_corrections_binary(meansq, mu, A, N, d)
    # a continuous annotation counts as its indicator
    ind = (A != 0).astype(float)
    Ma = ind.sum(0)
    S = ind.T @ meansq
    np.testing.assert_allclose(R2 * Ma[:, None] * Ma[None, :], S, rtol=1e-13)
    # with mu22 set to d^2 R2 / N the signal term vanishes: rho2 = 0, bias = S, delta = mu22 - 1
    mu0 = d * d * R2 / N
    R2b, rho2b, biasb, deltab = compute_block_corrections_binary(meansq, mu0, A, N, d)
    np.testing.assert_allclose(rho2b, 0.0, atol=1e-15)
    np.testing.assert_allclose(biasb, S, rtol=1e-13)
    np.testing.assert_allclose(deltab, mu0 - 1.0, rtol=1e-13)
    # an empty bin gives zeros in its row and column
    A2 = A.copy()
    A2[:, 1] = 0.0
    out = compute_block_corrections_binary(meansq, mu, A2, N, d)
    for arr in out:
        assert np.all(arr[1, :] == 0.0) and np.all(arr[:, 1] == 0.0)
    with pytest.raises(RuntimeError, match="mu22 shape mismatch"):
        compute_block_corrections_binary(meansq, mu[:2, :2], A, N, d)
    with pytest.raises(RuntimeError, match="N must be > 1"):
        compute_block_corrections_binary(meansq, mu, A, 1, d)
