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


@pytest.mark.parametrize("mode", ["mean", "hwe"])
def test_windowed_covariates_match_oracle(oracle, mode):
    """Covariate-projected, restandardized panels (winldcore.cpp:341-372) with a
    non-orthonormal C and R = (C^T C)^-1 C^T, a BP window and two annotation
    columns, against the oracle's dense projection."""
    print(f"windowed covariates: mode={mode}", flush=True)
    from summit_amd import engine, winldcore
    m, n = 520, 389
    codes = random_codes(m, n, seed=44, miss_rate=0.03, monomorphic=2)
    rng = np.random.default_rng(5)
    bp = np.cumsum(rng.integers(1000, 5000, size=m)).astype(np.int64)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.3).astype(float)])
    C = np.column_stack([np.ones(n), rng.standard_normal((n, 3)) * 2.0 + 0.5])
    R = np.linalg.solve(C.T @ C, C.T)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=0, impute=mode, impute_seed=13)
    lo, hi = 21, 500
    got = winldcore.windowed_ld(g, lo, hi, bp[lo:hi], 60.0, annot[lo:hi], C=C, R=R)
    g.close()
    ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, lo, hi, bp[lo:hi], annot[lo:hi], 60.0, impute=mode,
                                 impute_seed=13, C=C, R=R)
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-10)
    plain = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, lo, hi, bp[lo:hi], annot[lo:hi], 60.0, impute=mode,
                                   impute_seed=13)
    assert np.max(np.abs(plain - ref)) > 1e-3  # the covariates matter at this size


def test_windowed_covariates_dropin_orthonormal_basis(oracle):
    """The drop-in with the reference's own covariate form, C = Q and R = Q^T for
    Q = QR([1, covariates]) (gw_ldscore.py:1015-1035), row selection and a SNP
    slice; a bad C/R shape raises the reference's message."""
    from summit_amd import winldcore
    m, n = 260, 300
    codes = random_codes(m, n, seed=45, miss_rate=0.02)
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, "wc")
    oracle.write_bed(prefix, codes)
    rows = np.sort(np.random.default_rng(1).choice(n, 240, replace=False)).astype(np.int64)
    rng = np.random.default_rng(9)
    Q, _ = np.linalg.qr(np.column_stack([np.ones(240), rng.standard_normal((240, 2))]))
    C = np.asfortranarray(Q)
    R = np.asfortranarray(Q.T)
    bp = np.arange(m, dtype=np.int64) * 2000 + 7
    out = winldcore.compute_windowed_ld_chr(prefix, prefix + ".fam", 5, 250, bp[5:250], np.ones((245, 1)), 40.0, 50,
                                            row_sel=rows, C=C, R=R, impute_mode="mean")
    ref = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, 5, 250, bp[5:250], np.ones((245, 1)), 40.0, rows=rows,
                                 C=C, R=R)
    assert out.flags.f_contiguous
    np.testing.assert_allclose(out, ref, rtol=1e-10, atol=1e-10)
    with pytest.raises(RuntimeError, match="C/R shape mismatch"):
        winldcore.compute_windowed_ld_chr(prefix, prefix + ".fam", 5, 250, bp[5:250], np.ones((245, 1)), 40.0, 50,
                                          row_sel=rows, C=C[:-1], R=R, impute_mode="mean")


@pytest.mark.parametrize("mode", ["mean", "hwe"])
def test_one_sided_window_sums(oracle, mode):
    """Sums over the window SNPs before (side 1) and after (side 2) each SNP,
    against the oracle's masked unbiased r^2 matrix; with the SNP itself
    (r~2_jj = 1) the two sides give the two-sided sum."""
    from summit_amd import engine, winldcore
    m, n = 300, 333
    codes = random_codes(m, n, seed=46, miss_rate=0.02, monomorphic=1)
    annot = np.column_stack([np.ones(m), (np.arange(m) % 4 == 0).astype(float)])
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=0, impute=mode, impute_seed=3)
    lo, hi = 7, 290
    left = winldcore.windowed_ld(g, lo, hi, annot=annot[lo:hi], window_snps=40, side=1)
    right = winldcore.windowed_ld(g, lo, hi, annot=annot[lo:hi], window_snps=40, side=2)
    both = winldcore.windowed_ld(g, lo, hi, annot=annot[lo:hi], window_snps=40)
    g.close()
    mm = hi - lo
    R = oracle.windowed_ld_chr(oracle.pack_codes(codes), n, lo, hi, None, np.eye(mm), window_snps=40, impute=mode,
                               impute_seed=3)  # row j: r~2_ji within the window
    before = np.tril(R, -1) @ annot[lo:hi]
    after = np.triu(R, 1) @ annot[lo:hi]
    np.testing.assert_allclose(left, before, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(right, after, rtol=1e-11, atol=1e-11)
    np.testing.assert_allclose(left + right + annot[lo:hi] * np.diag(R)[:, None], both, rtol=1e-11, atol=1e-11)
