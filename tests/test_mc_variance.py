"""MC variance of the stochastic LD-score estimate from the probe sums
(src/summit/ldscore/gw_ldscore.py:874-995), checked against the sample variance
of explicit per-probe values."""
from __future__ import annotations

import numpy as np


def test_mc_variance_equals_sample_variance_over_v():
    from summit_amd.ldscore import integrated_mc_variance_from_probe_sums, mc_variance_from_probe_sums
    rng = np.random.default_rng(3)
    Y = rng.gamma(2.0, 1.5, size=(50, 3, 17))  # (snp, bin, probe)
    S1, S2 = Y.sum(axis=2), (Y * Y).sum(axis=2)
    got = mc_variance_from_probe_sums(S1, S2, 17)
    np.testing.assert_allclose(got, Y.var(axis=2, ddof=1) / 17, rtol=1e-12)
    agg = integrated_mc_variance_from_probe_sums(S1, S2.sum(axis=0), 17)
    np.testing.assert_allclose(agg, got.sum(axis=0), rtol=1e-11)


def test_mc_variance_roundoff_and_failures():
    from summit_amd.ldscore import mc_variance_from_probe_sums
    S1 = np.array([[3.0, 3.0, 3.0]])
    S2 = np.array([[3.0 - 1e-15, 1.0, np.inf]])  # tiny negative / large negative / non-finite
    res, info = mc_variance_from_probe_sums(S1, S2, 3, return_info=True)
    assert res[0, 0] == 0.0 and np.isnan(res[0, 1]) and np.isnan(res[0, 2])
    assert info["n_roundoff_clipped"].tolist() == [1, 0, 0]
    assert info["n_numerical_failures"].tolist() == [0, 1, 1]
    assert np.isnan(mc_variance_from_probe_sums(S1, S2, 1)).all()
