"""Pins of the oracle's GxE DirectContext restatement (direct_decode,
direct_features, direct_source, direct_target, direct_leakage), each by a route
other than its own formula, plus planted mistakes the checks reject."""
from __future__ import annotations

import numpy as np
import pytest


def _codes(rng, n, m, miss=0):
    g = rng.binomial(2, rng.uniform(0.15, 0.45, size=m)[:, None], size=(m, n)).astype(np.int64)
    if miss:
        g[rng.integers(0, m, size=miss), rng.integers(0, n, size=miss)] = -9
    return g


def _design(rng, n, ddof=1, n_cov=2):
    env = rng.normal(size=n)
    env = (env - env.mean()) / env.std(ddof=ddof)
    cov = rng.normal(size=(n, n_cov))
    q, _ = np.linalg.qr(np.column_stack([np.ones(n), env, cov]))
    return env, q


def _bed(oracle, g):
    return oracle.pack_codes(oracle.geno_to_codes(g))


@pytest.mark.parametrize("ddof", [0, 1])
def test_decode_matches_mean_filled_std_route(oracle, ddof):
    """The reference test's route (tests/test_gxe_native_core.py:95-105): the A1
    dosage 2 - g with missing calls filled by the column mean, centred and divided
    by std(ddof) over all N samples."""
    rng = np.random.default_rng(5)
    n, m = 41, 23
    g = _codes(rng, n, m, miss=9)
    X, missing = oracle.direct_decode(_bed(oracle, g), n, 0, m, ddof=ddof)
    raw = np.where(g < 0, np.nan, 2.0 - g).T
    means = np.nanmean(raw, axis=0)
    filled = np.where(np.isnan(raw), means[None, :], raw) - means[None, :]
    expect = filled / filled.std(axis=0, ddof=ddof)
    assert missing == int((g < 0).sum()) > 0
    assert np.max(np.abs(X - expect)) < 1e-13
    # planted mistake: the variance over the observed calls (nobs - ddof) differs
    # exactly where a SNP has missing calls
    codes, nobs, s, ss = oracle.decode_block(_bed(oracle, g), n, 0, m)
    wrong = np.sqrt((nobs - ddof) / (ss - s * s / nobs))
    right = np.sqrt((n - ddof) / (ss - s * s / nobs))
    assert np.any(np.abs(wrong - right) > 1e-6 * right)


def test_decode_is_negated_rescaled_genome_wide_decode(oracle):
    """x_direct = -x_gw sqrt((N - ddof) / (nobs - ddof)) per SNP: the relation the
    drop-in uses to serve DirectContext from the genome-wide standardization."""
    rng = np.random.default_rng(6)
    n, m = 37, 19
    g = _codes(rng, n, m, miss=12)
    g[3, :] = 1  # a monomorphic SNP: both forms are 0
    bed = _bed(oracle, g)
    Xd, _ = oracle.direct_decode(bed, n, 0, m, ddof=1)
    Xg = oracle.read_block_standardized(bed, n, 0, m, ddof=1, mode="mean")
    nobs = (g >= 0).sum(axis=1)
    r = np.sqrt((n - 1.0) / (nobs - 1.0))
    assert np.max(np.abs(Xd + Xg * r[None, :])) < 1e-13
    assert np.all(Xd[:, 3] == 0.0)


def _moment_features(X, env, q):
    """The reference's moment algebra (gxeldcore.cpp:8033-8243): scalar sums
    s_p = sum e^p x^2 and moments u_p = (E^p Q)^T x, no projected matrix formed."""
    n, p = q.shape
    rank = n - p
    x2 = X * X
    s0, s1, s2, s4 = (x2.sum(0), (env[:, None] * x2).sum(0), ((env ** 2)[:, None] * x2).sum(0),
                      ((env ** 4)[:, None] * x2).sum(0))
    u = [(q * (env ** k)[:, None]).T @ X for k in range(4)]
    qe2q = q.T @ ((env ** 2)[:, None] * q)
    u00, u11, u01 = (u[0] ** 2).sum(0), (u[1] ** 2).sum(0), (u[0] * u[1]).sum(0)
    u0u2, u1u3 = (u[0] * u[2]).sum(0), (u[1] * u[3]).sum(0)
    e2u0 = (u[0] * (qe2q @ u[0])).sum(0)
    e2u1 = (u[1] * (qe2q @ u[1])).sum(0)
    ssx, ssw = s0 - u00, s2 - u11
    sx, sw = np.sqrt(rank / ssx), np.sqrt(rank / ssw)
    return {"scale_x": sx, "scale_w": sw, "norm_x": sx * sx * ssx / rank, "norm_w": sw * sw * ssw / rank,
            "diag_nxe_x": sx * sx * (s2 - 2 * u0u2 + e2u0) / rank, "diag_nxe_w": sw * sw * (s4 - 2 * u1u3 + e2u1) / rank,
            "corr_xw": sx * sw * (s1 - u01) / rank}


def test_dense_features_match_moment_algebra(oracle):
    rng = np.random.default_rng(7)
    n, m = 53, 29
    g = _codes(rng, n, m, miss=5)
    env, q = _design(rng, n)
    X, _ = oracle.direct_decode(_bed(oracle, g), n, 0, m)
    dense = oracle.direct_features(X, env, q)
    moment = _moment_features(X, env, q)
    for k, v in moment.items():
        assert np.max(np.abs(dense[k] - v) / np.maximum(1.0, np.abs(v))) < 1e-12, k
    assert np.max(np.abs(dense["norm_x"] - 1.0)) < 1e-13
    assert np.max(np.abs(dense["norm_w"] - 1.0)) < 1e-13
    # the projected features are orthogonal to the design basis
    assert np.max(np.abs(q.T @ dense["x"])) < 1e-11
    assert np.max(np.abs(q.T @ dense["w"])) < 1e-11
    # planted mistake: the environment applied after the projection, W = e . (P x)
    xp = X - q @ (q.T @ X)
    w_bad = env[:, None] * xp
    corr_bad = np.sum(xp * w_bad, 0) / np.sqrt(np.sum(xp * xp, 0) * np.sum(w_bad * w_bad, 0))
    corr_good = moment["corr_xw"] / np.sqrt(moment["norm_x"] * moment["norm_w"])
    assert np.max(np.abs(corr_bad - corr_good)) > 1e-3


def test_source_and_target_brute_force(oracle):
    """Loops over every index on a tiny block."""
    rng = np.random.default_rng(8)
    n, m, V = 9, 4, 3
    g = _codes(rng, n, m, miss=2)
    env, q = _design(rng, n, n_cov=1)
    X, _ = oracle.direct_decode(_bed(oracle, g), n, 0, m)
    sx, sw = rng.uniform(0.5, 2.0, m), rng.uniform(0.5, 2.0, m)
    a = rng.uniform(0.0, 1.5, m)
    Z = rng.choice([-1.0, 1.0], size=(m, V))
    raw = np.zeros((n, 2 * V))
    for i in range(n):
        for c in range(V):
            for j in range(m):
                raw[i, c] += X[i, j] * a[j] * sx[j] * Z[j, c]
                raw[i, V + c] += env[i] * X[i, j] * a[j] * sw[j] * Z[j, c]
    proj = raw.copy()
    for c in range(2 * V):
        coef = [sum(q[i, k] * raw[i, c] for i in range(n)) for k in range(q.shape[1])]
        for i in range(n):
            proj[i, c] -= sum(q[i, k] * coef[k] for k in range(q.shape[1]))
    S = oracle.direct_source(X, env, q, sx, sw, a, Z)
    assert np.max(np.abs(S - proj)) < 1e-12
    assert oracle.direct_leakage(q, S) < 1e-14
    # targets of an arbitrary (unprojected) panel
    T = rng.normal(size=(n, 5))
    R = T - q @ (q.T @ T)
    wx, ww = oracle.direct_target(X, env, q, sx, sw, T)
    for j in range(m):
        for c in range(5):
            ex = sx[j] * sum(X[i, j] * R[i, c] for i in range(n))
            ew = sw[j] * sum(X[i, j] * env[i] * R[i, c] for i in range(n))
            assert abs(wx[j, c] - ex) < 1e-12 and abs(ww[j, c] - ew) < 1e-12
    # planted mistake: an unprojected panel gives other targets
    assert np.max(np.abs(wx - (X.T @ T) * sx[:, None])) > 1e-3


def test_leakage_closed_forms(oracle):
    rng = np.random.default_rng(9)
    n = 31
    env, q = _design(rng, n)
    assert abs(oracle.direct_leakage(q, q) - 1.0) < 1e-14
    S = rng.normal(size=(n, 4))
    S -= q @ (q.T @ S)
    assert oracle.direct_leakage(q, S) < 1e-15
    # a known fraction of the basis mixed in: leakage = ||c|| / ||S + q c||
    c = np.array([[0.3, 0.0, 0.0, 0.0]])
    mixed = S + q[:, :1] @ c
    expect = 0.3 / np.sqrt(np.sum(mixed * mixed))
    assert abs(oracle.direct_leakage(q, mixed) - expect) < 1e-14
