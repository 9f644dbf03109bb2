"""Pins of the CPU oracle (runs without a GPU).

Each function of oracle/oracle.py is checked against something other than a
retyping of itself: golden vectors written by the reference's own code
(tests/golden/make_golden.py), known answers, brute force on tiny inputs,
and the statistical test of the reference's own suite.
"""
from __future__ import annotations

import ctypes
import os
import tempfile

import numpy as np
import pytest

from tests.helpers import random_codes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libref_genotype.so")


def test_mt19937_64_known_answer(oracle):
    # C++ standard [rand.predef]: 10000th output of a default-constructed mt19937_64
    assert oracle.mt64_nth(5489, 10000) == 9981545732273789042


def test_make_seed_mixes_every_argument(oracle):
    base = oracle.make_seed(1, 0, 0)
    assert len({base, oracle.make_seed(2, 0, 0), oracle.make_seed(1, 1, 0), oracle.make_seed(1, 0, 1)}) == 4


def test_probe_streams_match_reference_golden(oracle):
    """Rademacher, Gaussian and spherical streams in double and float precision against
    libstdc++'s own std::mt19937_64 / std::normal_distribution<T> driven by the
    reference's fill (vectors written by tests/golden/make_golden.py)."""
    g = np.load(os.path.join(GOLDEN, "probe_streams.npz"))
    assert len(g.files) == 15
    for key in g.files:
        dist, root, blk, v0, L, V = key.split("_")
        fp32 = dist.endswith("32")
        Z = oracle.probe_block(int(root), int(blk), int(v0), int(L), int(V), dist.rstrip("32"), fp32=fp32)
        np.testing.assert_array_equal(Z, g[key], err_msg=key)


def test_spherical_columns_have_norm_sqrt_l(oracle):
    """Spherical probes: each column of a block has squared norm L (to rounding)."""
    for fp32 in (False, True):
        Z = oracle.probe_block(9, 3000, 5, 41, 6, "spherical", fp32=fp32)
        np.testing.assert_allclose((Z * Z).sum(axis=0), 41.0, rtol=1e-5 if fp32 else 1e-13)


def test_rademacher_uses_low_bit_column_major(oracle):
    raw = oracle.stream_raw(77, 5, 9, 12)
    Z = oracle.probe_block(77, 5, 9, 4, 3)
    expect = np.where(raw & np.uint64(1), 1.0, -1.0).reshape(3, 4).T
    np.testing.assert_array_equal(Z, expect)


def test_bed_byte_layout_known_answer(oracle):
    # PLINK: sample 4b+t sits in bits 2t..2t+1; 0b11_10_01_00 = codes 0,1,2,3
    codes = np.array([[0, 1, 2, 3, 3]], dtype=np.uint8)
    packed = oracle.pack_codes(codes)
    assert packed.tolist() == [[0b11100100, 0b00000011]]
    np.testing.assert_array_equal(oracle.unpack_codes(packed, 5), codes)
    np.testing.assert_array_equal(oracle.codes_to_geno(codes), [[0, -9, 1, 2, 2]])


def test_bed_file_roundtrip_and_magic(oracle):
    codes = random_codes(9, 13, seed=1)
    d = tempfile.mkdtemp()
    p = os.path.join(d, "x")
    oracle.write_bed(p, codes)
    np.testing.assert_array_equal(oracle.unpack_codes(oracle.read_bed_matrix(p, 13, 9), 13), codes)
    with open(p + ".bed", "r+b") as f:
        f.write(b"\x6c\x1b\x00")
    with pytest.raises(RuntimeError, match="magic"):
        oracle.read_bed_matrix(p, 13, 9)


@pytest.mark.parametrize("mode", ["mean", "hwe"])
@pytest.mark.parametrize("subset", [False, True])
def test_standardize_matches_reference_golden(oracle, mode, subset):
    g = np.load(os.path.join(GOLDEN, "genotype_decode.npz"))
    codes = g["codes"]
    bed = oracle.pack_codes(codes)
    rows = g["rows"] if subset else None
    X = oracle.read_block_standardized(bed, codes.shape[1], int(g["blk_start"]), int(g["blk_end"]), rows, 1, mode,
                                       int(g["impute_seed"]))
    np.testing.assert_array_equal(X, g[f"X_{mode}_{'sub' if subset else 'all'}"])


def test_hwe_stats_match_reference_mailman_pack(oracle):
    g = np.load(os.path.join(GOLDEN, "genotype_decode.npz"))
    codes, rows = g["codes"], g["rows"]
    s0, s1 = int(g["blk_start"]), int(g["blk_end"])
    c, nobs, s, ss = oracle.decode_block(oracle.pack_codes(codes), codes.shape[1], s0, s1, rows)
    _, mu, inv = oracle.standardize_block(c, nobs, s, ss, 1, "hwe", int(g["impute_seed"]), s0, rows)
    np.testing.assert_array_equal(mu, g["hwe_mailman_mean_sub"])
    np.testing.assert_array_equal(inv, g["hwe_mailman_inv_std_sub"])


def test_standardize_matches_textbook_column_standardization(oracle):
    # tests/test_pgen_gwld.py:244-262: observed-only mean/sd, missing -> 0
    codes = random_codes(25, 60, seed=9, miss_rate=0.1, monomorphic=2, all_missing=1)
    X = oracle.read_block_standardized(oracle.pack_codes(codes), 60, 0, 25)
    raw = oracle.codes_to_geno(codes).astype(np.float64).T
    for j in range(25):
        col = raw[:, j]
        obs = col != -9
        expect = np.zeros(60)
        if obs.any():
            vals = col[obs]
            cen = vals - vals.mean()
            inv_sd = 1.0
            if vals.size - 1 > 0 and cen @ cen > 0:
                inv_sd = 1.0 / np.sqrt((cen @ cen) / (vals.size - 1))
            expect[obs] = cen * inv_sd
        np.testing.assert_allclose(X[:, j], expect, rtol=1e-13, atol=1e-13)


@pytest.mark.skipif(not os.path.exists(REF_LIB), reason="reference tree not built here (oracle/_ref)")
@pytest.mark.parametrize("mode,mi", [("mean", 0), ("hwe", 1)])
def test_standardize_matches_reference_build_live(oracle, mode, mi):
    ref = ctypes.CDLL(REF_LIB)
    ref.ref_read_block_standardized_double.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    codes = random_codes(150, 257, seed=4, miss_rate=0.05, monomorphic=3, all_missing=2)
    d = tempfile.mkdtemp()
    p = os.path.join(d, "t")
    oracle.write_bed(p, codes)
    rows = np.sort(np.random.default_rng(2).choice(257, 190, replace=False)).astype(np.int32)
    out = np.empty((100, rows.size))
    assert ref.ref_read_block_standardized_double((p + ".bed").encode(), (p + ".fam").encode(), 25, 125,
                                                  rows.ctypes.data, rows.size, 1, mi, 987654321, out.ctypes.data) == 0
    X = oracle.read_block_standardized(oracle.pack_codes(codes), 257, 25, 125, rows, 1, mode, 987654321)
    np.testing.assert_array_equal(X, out.T)


def test_phase2_dense_matrix_oracle(oracle):
    # inputs of tests/test_gwld_mc.py:42-128, expected values by explicit loops
    geno = np.array([[0.2, -0.4, 0.8, 0.1], [1.1, 0.3, -0.2, 0.5], [-0.7, 0.9, 0.4, -0.6],
                     [0.6, -1.2, 0.7, 0.3], [0.1, 0.8, -0.5, 1.0]])
    xz = np.array([[0.3, -0.1, 0.7, 0.2, 0.5, -0.4], [1.0, 0.6, -0.2, 0.8, -0.3, 0.9],
                   [-0.5, 0.4, 0.1, -0.7, 0.2, 0.3], [0.9, -0.8, 0.3, 0.4, 0.7, -0.1],
                   [0.2, 0.5, -0.6, 1.1, -0.9, 0.6]])
    inv = np.array([1.0, 0.75, 1.25, 0.6])
    bins, probes, n_denom = 2, 3, 6
    point = np.zeros((4, bins))
    agg = np.zeros(bins)
    per = np.zeros((4, bins))
    oracle.phase2_accum(np.asfortranarray(geno), inv, probes, np.asfortranarray(xz), point, n_denom, 0, agg, per)
    exp = np.zeros((4, bins))
    exp4 = np.zeros((4, bins))
    for j in range(4):
        for k in range(bins):
            for v in range(probes):
                w = sum(geno[n, j] * xz[n, k * probes + v] for n in range(5))
                y = w * w * inv[j] ** 2 / (n_denom - 1) ** 2
                exp[j, k] += y
                exp4[j, k] += y * y
    np.testing.assert_allclose(point, exp, rtol=2e-14)
    np.testing.assert_allclose(per, exp4, rtol=2e-14)
    np.testing.assert_allclose(agg, exp4.sum(axis=0), rtol=2e-14)


def test_phase2_guards_drop_nonfinite_and_extreme(oracle):
    X = np.asfortranarray(np.array([[1.0, 1e9], [1.0, -1e9], [0.0, 3e9]]))
    xz = np.asfortranarray(np.array([[1.0], [2.0], [np.nan]]))
    out = np.zeros((2, 1))
    counts = oracle.phase2_accum(X, np.array([1.0, 1.0]), 1, xz, out, 3, 0)
    assert counts["nonfinite"] == 2 and out.sum() == 0.0
    xz2 = np.asfortranarray(np.array([[1.0], [2.0], [3.0]]))
    out2 = np.zeros((2, 1))
    counts2 = oracle.phase2_accum(X, np.array([1.0, 1.0]), 1, xz2, out2, 3, 0)
    assert counts2["extreme"] == 1 and out2[1, 0] == 0.0 and out2[0, 0] > 0.0


def test_phase1_bin_scale_skip_rules(oracle):
    annot = np.array([[1.0, 0.0, -1.0, np.inf, 4.0]])
    inv = np.array([2.0])
    np.testing.assert_array_equal(oracle.bin_scales(annot, inv), [[2.0, 0.0, 0.0, 0.0, 4.0]])
    assert oracle.bin_scales(np.array([[1.0]]), np.array([0.0]))[0, 0] == 0.0
    assert oracle.bin_scales(np.array([[1.0]]), np.array([np.nan]))[0, 0] == 0.0


def test_integer_products_brute_force(oracle):
    rng = np.random.default_rng(0)
    codes = random_codes(7, 11, seed=3, miss_rate=0.2)
    g = oracle.masked_geno(codes)
    z = rng.integers(-3, 4, size=(7, 4))
    u = oracle.gz_int(g, z)
    for n in range(11):
        for b in range(4):
            assert u[n, b] == sum(int(g[i, n]) * int(z[i, b]) for i in range(7))
    w = oracle.gtu_int(g, u)
    for i in range(7):
        for b in range(4):
            assert w[i, b] == sum(int(g[i, n]) * int(u[n, b]) for n in range(11))


def test_vtile_sizes(oracle):
    # widths >= 64 are rounded down to a multiple of 64 (gw_ldscore.py:694-697)
    assert oracle.bounded_vtile_sizes(100, 1000) == [64, 36]
    assert oracle.bounded_vtile_sizes(64, 1000) == [64]
    assert oracle.bounded_vtile_sizes(1000, 300) == [256, 256, 256, 232]
    assert oracle.bounded_vtile_sizes(50, 10) == [10] * 5


def test_estimator_within_five_monte_carlo_se_of_exact_ld(oracle):
    """Statistical pin of tests/test_pgen_gwld.py:715-772: |L_hat - L| <= 5 SE with
    probe variance 2[(sum r^2)^2 - sum r^4] / B."""
    codes = random_codes(8, 40, seed=12, miss_rate=0.05)
    bed = oracle.pack_codes(codes)
    nvecs = 20000
    res = oracle.gw_ldscore(bed, 40, 8, nvecs=nvecs, step_size=3, seed=409, vtiles=[8000, 8000, 4000])
    exact, corr = oracle.exact_ld(bed, 40, 8)
    r2 = corr * corr
    var = 2.0 * (np.square(r2.sum(axis=1)) - np.sum(r2 * r2, axis=1))
    se = np.sqrt(np.maximum(var, 0.0) / nvecs)
    err = np.abs(res.ldscore[:, 0] - exact[:, 0])
    assert np.all(err <= 5.0 * se + 1e-10), (err, se)


def test_partitioned_estimator_targets_annotated_sums(oracle):
    codes = random_codes(10, 50, seed=13, miss_rate=0.02)
    bed = oracle.pack_codes(codes)
    annot = np.zeros((10, 2))
    annot[:6, 0] = 1.0
    annot[4:, 1] = 1.0
    nvecs = 20000
    res = oracle.gw_ldscore(bed, 50, 10, annot=annot, nvecs=nvecs, step_size=4, seed=7, vtiles=[10000, 10000])
    exact, corr = oracle.exact_ld(bed, 50, 10, annot=annot)
    r2 = corr * corr
    for k in range(2):
        rk = r2 * annot[None, :, k]
        var = 2.0 * (np.square(rk.sum(axis=1)) - np.sum(rk * rk, axis=1))
        se = np.sqrt(np.maximum(var, 0.0) / nvecs)
        assert np.all(np.abs(res.ldscore[:, k] - exact[:, k]) <= 5.0 * se + 1e-10)
