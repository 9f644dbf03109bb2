"""Phase-2 update guards on the device (src/native/gwldcore.cpp:295-417): a
non-finite update is counted and dropped, an update above 1e10 x V is counted
and dropped (or kept when dropping is off); counts and accumulators against
the oracle's restatement fed the same inputs."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu


def _setup(oracle, m, n, seed):
    import torch

    from summit_amd import engine
    codes = random_codes(m, n, seed=seed, miss_rate=0.02)
    bed = oracle.pack_codes(codes)
    g = engine.Genotype.from_bed_bytes(bed, n, m).prepare(impute="mean")
    X = oracle.read_block_standardized(bed, n, 0, m, None, 1, "mean", 0)
    return g, np.asfortranarray(X), torch


@pytest.mark.parametrize("drop_extreme", [True, False])
def test_guards_nonfinite_column_and_extreme_snp(oracle, drop_extreme):
    print(f"p2 guards: drop_extreme={drop_extreme}", flush=True)
    from summit_amd import engine
    m, n, V = 300, 200, 4
    g, X, torch = _setup(oracle, m, n, 41)
    rng = np.random.default_rng(2)
    xz = np.asfortranarray(rng.standard_normal((n, 3 * V)))
    xz[17, V + 2] = np.nan          # bin 1: one NaN in one column
    xz[5, 2 * V + 1] = np.inf       # bin 2: one infinity
    inv = rng.uniform(0.9, 1.1, size=m)
    inv[[7, 150]] = 1e8             # SNPs whose update exceeds 1e10 x V in bin 0
    inv[[20, 21]] = 0.0             # unscaled SNPs (left scale 0)
    ref = np.full((m, 3), 0.25)
    mc_ref = np.zeros((m, 3))
    counts = oracle.phase2_accum(X, inv, V, xz, ref, n, 0, mc_per_snp_sumsq=mc_ref, drop_extreme=drop_extreme)
    dev = engine.default_device()
    ms = torch.full((m, 3), 0.25, dtype=torch.float64, device=dev)
    mcp = torch.zeros((m, 3), dtype=torch.float64, device=dev)
    mca = torch.zeros(3, dtype=torch.float64, device=dev)
    got = engine.phase2_accum(g, 0, m, torch.from_numpy(inv).to(dev), V,
                              torch.from_numpy(np.ascontiguousarray(xz.T)).to(dev), ms, n,
                              mc_aggregate_sumsq=mca, mc_per_snp_sumsq=mcp, extreme_warn=1e10,
                              drop_extreme=drop_extreme, limbs=5, want_guards=True)
    g.close()
    # every SNP of bins 1 and 2 is non-finite (the unscaled ones too: NaN x 0);
    # SNPs 7 and 150 are extreme in bin 0
    assert counts == {"nonfinite": 2 * m, "extreme": 2}
    assert got == (counts["nonfinite"], counts["extreme"])
    out = ms.cpu().numpy()
    assert np.all(out[:, 1:] == 0.25)  # dropped updates leave the accumulators alone
    if drop_extreme:
        assert out[7, 0] == 0.25 and out[150, 0] == 0.25
    else:
        assert out[7, 0] > 1e10 * V
    np.testing.assert_allclose(out, ref, rtol=5e-11)
    np.testing.assert_allclose(mcp.cpu().numpy(), mc_ref, rtol=2e-10)
    np.testing.assert_allclose(mca.cpu().numpy(), mc_ref.sum(axis=0), rtol=2e-10)


def test_dropin_reports_guards_and_keeps_finite_bins(oracle, capfd):
    """Through the drop-in: the guard message names the block and the counts,
    and the finite bin still matches the oracle."""
    import os
    import tempfile

    from summit_amd import gwldcore as gw
    m, n, V = 120, 150, 3
    codes = random_codes(m, n, seed=43, miss_rate=0.02)
    prefix = os.path.join(tempfile.mkdtemp(), "gd")
    oracle.write_bed(prefix, codes)
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, "mean", 0)
    rng = np.random.default_rng(3)
    xz = np.asfortranarray(rng.standard_normal((n, 2 * V)))
    xz[:, V] = np.nan
    inv = rng.uniform(0.9, 1.1, size=m)
    meansq = np.zeros((m, 2))
    gw.phase2_accum_XtXz_bed(bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=0, blk_end=m, inv_left=inv,
                             tile_nvecs=V, Xz2d=xz, meansq_accum=meansq, N_denom=n, impute_mode="mean")
    err = capfd.readouterr().err
    assert "[phase2-watch] accum block [0:120) nonfinite_dropped=120 extreme=0" in err
    ref = np.zeros((m, 2))
    oracle.phase2_accum(np.asfortranarray(X), inv, V, xz, ref, n, 0)
    assert np.all(meansq[:, 1] == 0.0)
    np.testing.assert_allclose(meansq, ref, rtol=5e-11)
