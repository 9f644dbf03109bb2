"""Exact windowed LD scores on the GPU against the oracle restatement of winldcore."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("mode", ["mean", "hwe"])
def test_windowed_bp_window_matches_oracle(oracle, mode):
    from summit_amd import engine, winldcore
    m, n = 700, 413
    codes = random_codes(m, n, seed=41, miss_rate=0.02, monomorphic=2)
    bp = np.cumsum(np.random.default_rng(2).integers(1000, 5000, size=m)).astype(np.int64)
    annot = np.column_stack([np.ones(m), (np.arange(m) % 3 == 0).astype(float)])
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=0, impute=mode, impute_seed=11)
    lo, hi = 37, 650
    got = winldcore.windowed_ld(g, lo, hi, bp[lo:hi], 100.0, annot[lo:hi])
    ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, lo, hi, bp[lo:hi], annot[lo:hi], 100.0,
                                 impute=mode, impute_seed=11)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11)
    g.close()


def test_windowed_snp_window_and_dropin(oracle):
    from summit_amd import engine, winldcore
    m, n = 300, 257
    codes = random_codes(m, n, seed=42, miss_rate=0.01)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=0, impute="mean")
    got = winldcore.windowed_ld(g, 0, m, window_snps=50)
    ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, 0, m, None, np.ones((m, 1)), window_snps=50)
    np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11)
    g.close()
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, "w")
    oracle.write_bed(prefix, codes)
    bp = np.arange(m, dtype=np.int64) * 1000 + 5
    out = winldcore.compute_windowed_ld_chr(prefix, prefix + ".fam", 10, 290, bp[10:290], np.ones((280, 1)), 20.0,
                                            50, impute_mode="mean")
    ref2 = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, 10, 290, bp[10:290], np.ones((280, 1)), 20.0)
    assert out.flags.f_contiguous
    np.testing.assert_allclose(out, ref2, rtol=1e-11, atol=1e-11)
    with pytest.raises(RuntimeError, match="BP not sorted"):
        winldcore.compute_windowed_ld_chr(prefix, prefix + ".fam", 0, 5, np.array([5, 4, 3, 2, 1]),
                                          np.ones((5, 1)), 1.0, 10)


@pytest.mark.parametrize("mode", ["mean", "hwe"])
def test_windowed_all_chromosomes_in_one_launch(oracle, mode):
    """Three chromosomes of uneven length (one shorter than a tile) and a gap
    between two of them in one launch: each chromosome's scores equal the
    oracle's per-chromosome windowed LD; rows in the gap stay 0."""
    from summit_amd import engine, winldcore
    m, n = 900, 311
    codes = random_codes(m, n, seed=43, miss_rate=0.02, monomorphic=1)
    rng = np.random.default_rng(5)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=0, impute=mode, impute_seed=3)
    chroms = []
    for s0, s1 in ((10, 420), (420, 457), (500, 890)):
        bp = np.cumsum(rng.integers(500, 4000, size=s1 - s0)).astype(np.int64)
        chroms.append((s0, s1, bp))
    annot = np.column_stack([np.ones(880), (np.arange(880) % 4 == 1).astype(float)])
    got = winldcore.windowed_ld_multi(g, chroms, ld_wind_kb=60.0, annot=annot)
    assert got.shape == (880, 2)
    for s0, s1, bp in chroms:
        ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, s0, s1, bp, annot[s0 - 10:s1 - 10], 60.0,
                                     impute=mode, impute_seed=3)
        np.testing.assert_allclose(got[s0 - 10:s1 - 10], ref, rtol=1e-11, atol=1e-11)
        one = winldcore.windowed_ld(g, s0, s1, bp, 60.0, annot[s0 - 10:s1 - 10])
        np.testing.assert_allclose(got[s0 - 10:s1 - 10], one, rtol=1e-12, atol=1e-12)
    assert np.all(got[457 - 10:500 - 10] == 0.0)
    snp = winldcore.windowed_ld_multi(g, [(c[0], c[1], None) for c in chroms], annot=annot, window_snps=25)
    for s0, s1, _ in chroms:
        ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, s0, s1, None, annot[s0 - 10:s1 - 10],
                                     window_snps=25, impute=mode, impute_seed=3)
        np.testing.assert_allclose(snp[s0 - 10:s1 - 10], ref, rtol=1e-11, atol=1e-11)
    with pytest.raises(RuntimeError, match="sorted and must not overlap"):
        winldcore.windowed_ld_multi(g, [(100, 200, None), (150, 250, None)], window_snps=5)
    g.close()
