"""The K-moment summary and probe draws (summit_amd.kmoments) against closed
forms of known spectra (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from summit_amd import kmoments


def _exact_row(lam, r, q=7):
    """The summary of exact traces sum lambda^p, every probe giving the exact value."""
    tr = [float(np.sum(lam ** p)) for p in (1, 2, 3, 4)]
    ks = [np.full(q, t) for t in tr]
    return kmoments.summarize(*ks, tr[1], r, 1000, 5)


def test_summary_measures_departure_from_a_projection():
    rng = np.random.default_rng(0)
    r = 40
    lam = rng.uniform(0.2, 2.5, size=r)
    lam *= r / lam.sum()  # sum lambda = r, so tr(K) = r and alpha = 1
    row = _exact_row(lam, r)
    d = lam - 1.0
    assert row["alpha_probe"] == pytest.approx(1.0, rel=1e-14)
    assert row["t0_rank"] == pytest.approx(np.sum(d ** 2), rel=1e-12)
    assert row["t1_rank"] == pytest.approx(np.sum(lam * d ** 2), rel=1e-12)
    assert row["t2_rank"] == pytest.approx(np.sum(lam ** 2 * d ** 2), rel=1e-12)
    assert row["delta_reff_rank"] == pytest.approx(np.sum(d ** 2) ** 2 / np.sum(d ** 4), rel=1e-10)
    # with alpha = 1 the probe-alpha forms coincide with the rank forms
    for k in ("t0", "t1", "t2"):
        assert row[f"{k}_probealpha"] == pytest.approx(row[f"{k}_rank"], rel=1e-10, abs=1e-9)
    assert row["trace_K_probe_se"] == 0.0 and row["alpha_flag"] == "OK" and row["mc_flag"] == "OK"
    # a spike: one eigenvalue far from 1 -> effective rank near 1
    lam2 = np.ones(r)
    lam2[0] += 20.0
    lam2[1:] -= 20.0 / (r - 1)
    assert _exact_row(lam2, r)["delta_reff_rank"] < 2.0
    assert _exact_row(lam2, r)["spike_flag"] == "LOW_EFFECTIVE_RANK"


def test_summary_of_an_exact_projection():
    row = _exact_row(np.ones(25), 25)
    assert row["t0_rank"] == 0.0 and row["t1_rank"] == 0.0 and row["t2_rank"] == 0.0
    assert np.isnan(row["delta_reff_rank"]) and row["spike_flag"] == "NA"


def test_alpha_scaling_and_flags():
    # K = c P (a rank-r projection scaled by c): alpha = c; the probe-alpha forms
    # remove the scale (t0 = c^2 r - c^2 r = 0 for the alpha form)
    r, c = 30, 1.02
    lam = np.full(r, c)
    row = _exact_row(lam, r)
    assert row["alpha_probe"] == pytest.approx(c, rel=1e-14)
    assert row["t0_probealpha"] == pytest.approx(0.0, abs=1e-9)
    assert row["t2_probealpha"] == pytest.approx(0.0, abs=1e-9)
    assert row["alpha_flag"] == "BAD"


def test_probe_draws():
    assert kmoments.probe_seed(0) == 0xD1B54A32D192ED03
    assert kmoments.probe_seed(0xD1B54A32D192ED03) == 0
    z = kmoments.sample_probes(200, 16, "rademacher", 7)
    assert z.flags.f_contiguous and set(np.unique(z)) == {-1.0, 1.0}
    np.testing.assert_array_equal(z, kmoments.sample_probes(200, 16, "rademacher", 7))
    s = kmoments.sample_probes(200, 16, "spherical", 7)
    np.testing.assert_allclose(np.linalg.norm(s, axis=0), np.sqrt(200), rtol=1e-13)
    # the Gaussian draws are NumPy's standard normals of the seeded generator, in order
    gz = kmoments.sample_probes(50, 3, "gaussian", 11)
    np.testing.assert_array_equal(gz, np.random.default_rng(11).standard_normal((50, 3)))
    with pytest.raises(ValueError, match="Unsupported rand_dist"):
        kmoments.sample_probes(5, 2, "uniform", 1)


def test_trace_k2_from_ldscore_and_writer(tmp_path):
    # meansq_j = ||X^T x_j||^2 / r^2 summed over j is ||X^T X||_F^2 / r^2 = tr(K^2) M^2 / r^2
    assert kmoments.trace_k2_from_ldscore(np.full(10, 2.0), 10, 5.0) == pytest.approx(25.0 * 20.0 / 100.0)
    row = _exact_row(np.ones(4), 4, q=3)
    path = kmoments.write_gw_kmoments(str(tmp_path / "o"), row)
    lines = open(path).read().splitlines()
    assert lines[0].split("\t")[:4] == ["nsnps", "proj_rank", "num_probes", "probe_seed"]
    assert lines[1].split("\t")[:3] == ["1000", "4.0000000000e+00", "3"]
