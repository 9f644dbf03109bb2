"""GPU parity tests of libsummit_gpe against the CPU oracle (bit-exact for
integer/byte work, FP64 tolerances stated per test for the estimator)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from tests.helpers import assert_ld_close, random_codes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg: str) -> None:
    print(f"[gpu-case] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def engine():
    from summit_amd import engine as e
    e.default_device()
    return e


def _geno(engine, oracle, codes, rows=None, n_total=None):
    m, n = codes.shape
    body = oracle.pack_codes(codes)
    return engine.Genotype.from_bed_bytes(body, n if n_total is None else n_total, m, rows=rows)


@pytest.mark.parametrize("m,n", [(1, 2), (37, 129), (300, 1000), (1000, 517), (129, 4097)])
def test_upload_roundtrip_and_layouts(engine, oracle, m, n):
    _say(f"layouts m={m} n={n}")
    codes = random_codes(m, n, seed=m * 7 + n)
    g = _geno(engine, oracle, codes)
    np.testing.assert_array_equal(g.download("snp_raw"), oracle.pack_codes(codes))
    g.prepare(ddof=1, impute="mean")
    np.testing.assert_array_equal(g.download("sample_work"), oracle.pack_codes(codes.T))
    st = g.stats()
    _, nobs, s, ss = oracle.decode_block(oracle.pack_codes(codes), n, 0, m)
    np.testing.assert_array_equal(st["nobs"], nobs)
    np.testing.assert_array_equal(st["sum"], s)
    np.testing.assert_array_equal(st["sumsq"], ss)
    mu, inv = oracle._mean_invsd(nobs, s, ss, 1)
    np.testing.assert_allclose(st["mean"], mu, rtol=1e-15, atol=0)
    np.testing.assert_allclose(st["inv_sd"], inv, rtol=1e-15, atol=0)
    assert g.n_missing() == int((codes == 1).sum())
    g.close()


def test_row_selection_and_hwe_completion(engine, oracle):
    _say("row_sel + hwe")
    m, n_total = 211, 901
    codes = random_codes(m, n_total, seed=5, miss_rate=0.05, monomorphic=3, all_missing=2)
    rows = np.sort(np.random.default_rng(1).choice(n_total, size=640, replace=False))
    g = _geno(engine, oracle, codes, rows=rows, n_total=n_total)
    np.testing.assert_array_equal(g.download("snp_raw"), oracle.pack_codes(codes[:, rows]))
    seed = 0xA24BAED4963EE407 ^ 1234
    g.prepare(ddof=1, impute="hwe", impute_seed=seed, snp_offset=1000, global_rows=rows)
    sub = codes[:, rows]
    obs = sub != 1
    val = np.array([0, 0, 1, 2])[sub]
    comp, s2, ss2 = oracle.hwe_impute(sub, obs.sum(1), val.sum(1), seed, 1000, rows)
    np.testing.assert_array_equal(g.download("snp_work"), oracle.pack_codes(comp))
    np.testing.assert_array_equal(g.download("sample_work"), oracle.pack_codes(comp.T))
    st = g.stats()
    np.testing.assert_array_equal(st["sum"], s2)
    mu, inv = oracle._mean_invsd(np.full(m, len(rows)), s2, ss2, 1)
    np.testing.assert_allclose(st["mean"], mu, rtol=1e-15, atol=0)
    np.testing.assert_allclose(st["inv_sd"], inv, rtol=1e-15, atol=0)
    assert g.n_missing() == 0
    g.close()


@pytest.mark.parametrize("m,n,b", [(37, 129, 3), (300, 1000, 64), (1000, 517, 10), (2049, 300, 100)])
def test_exact_integer_products(engine, oracle, m, n, b):
    _say(f"exact gz/gtu m={m} n={n} b={b}")
    codes = random_codes(m, n, seed=m + n + b, miss_rate=0.03)
    g = _geno(engine, oracle, codes).prepare()
    rng = np.random.default_rng(b)
    z = rng.choice(np.array([-1, 1], dtype=np.int8), size=(m, b))
    zt = torch.from_numpy(z).cuda()
    u = engine.gz_i32(g, zt)
    gm = oracle.masked_geno(codes)
    u_ref = oracle.gz_int(gm, z)  # (n, b)
    np.testing.assert_array_equal(u.cpu().numpy().T, u_ref)
    w = engine.gtu_i64(g, u)
    np.testing.assert_array_equal(w.cpu().numpy(), oracle.gtu_int(gm, u_ref))
    # general int8 coefficients and a sub-range of SNPs, accumulate
    z2 = rng.integers(-128, 128, size=(m // 2 + 1, b), dtype=np.int8)
    lo = m // 3
    hi = lo + z2.shape[0] if lo + z2.shape[0] <= m else m
    z2 = z2[: hi - lo]
    u2 = engine.gz_i32(g, torch.from_numpy(z2).cuda(), lo, hi, out=u.clone(), accumulate=True)
    np.testing.assert_array_equal(u2.cpu().numpy().T, u_ref + oracle.gz_int(gm[lo:hi], z2))
    w2 = engine.gtu_i64(g, u, lo, hi)
    np.testing.assert_array_equal(w2.cpu().numpy(), oracle.gtu_int(gm[lo:hi], u_ref))
    g.close()


@pytest.mark.parametrize("dist,fp32", [("rademacher", False), ("gaussian", False), ("spherical", False),
                                       ("gaussian", True), ("spherical", True)])
def test_probe_streams(engine, oracle, dist, fp32):
    _say(f"probes {dist} fp32={fp32}")
    starts = [0, 1000, 2000, 2500]
    end = 3333
    z = engine.probes(1101, starts, end, v_start=64, v_count=13, rand_dist=dist, fp32=fp32)
    ref = np.concatenate([oracle.probe_block(1101, s, 64, e - s, 13, dist, fp32=fp32)
                          for s, e in zip(starts, starts[1:] + [end])])
    np.testing.assert_array_equal(z, ref)


def test_rademacher_long_block_streams(engine, oracle):
    """Long and short SNP blocks in one launch (13,000 / 500 / 500 SNPs x 40
    probes): each stream's output index is advanced without divisions."""
    _say("probes rademacher long blocks")
    starts = [0, 13000, 13500]
    end = 14000
    z = engine.probes(1101, starts, end, v_start=64, v_count=40, rand_dist="rademacher")
    ref = np.concatenate([oracle.probe_block(1101, s, 64, e - s, 40, "rademacher")
                          for s, e in zip(starts, starts[1:] + [end])])
    np.testing.assert_array_equal(z, ref)


@pytest.mark.parametrize("fp32", [False, True])
def test_phase1_gaussian_probes_many_blocks_match_oracle(engine, oracle, fp32):
    """Phase 1 with Gaussian probes over enough SNP blocks that the host draws
    are cut into parts (each part's traversal queued while the next is drawn):
    the first part writes the set-mode bins, later parts add."""
    _say(f"phase1 gaussian many blocks fp32={fp32}")
    m, n, K, V, step = 2100, 300, 2, 7, 100
    codes = random_codes(m, n, seed=17, miss_rate=0.02)
    rng = np.random.default_rng(8)
    annot = np.ones((m, K))
    annot[:, 1] = rng.random(m) < 0.4
    inv = rng.uniform(0.8, 1.2, size=m)
    g = _geno(engine, oracle, codes).prepare(impute="mean")
    prior = rng.standard_normal((K * V, n))
    xz = torch.from_numpy(prior.copy()).cuda()
    mode = torch.tensor([1, 2], dtype=torch.uint8, device="cuda")  # bin 0 set, bin 1 add
    engine.phase1_xz(g, 0, m, step, torch.from_numpy(annot).cuda(), torch.from_numpy(inv).cuda(), 5, V, 77, xz,
                     rand_dist="gaussian", bin_mode=mode, probe_fp32=fp32)
    bed = oracle.pack_codes(codes)
    ref = np.zeros((n, K * V))
    for s in range(0, m, step):
        e = min(m, s + step)
        X = oracle.read_block_standardized(bed, n, s, e, None, 1, "mean", 0)
        ref += oracle.phase1_chunk(X, annot[s:e], inv[s:e], oracle.probe_block(77, s, 5, e - s, V, "gaussian",
                                                                                fp32=fp32))
    ref[:, V:] += prior[V:].T
    err = np.abs(xz.cpu().numpy().T - ref) / np.abs(ref).max(axis=0, keepdims=True)
    assert err.max() < 3e-11, err.max()
    g.close()


def test_phase1_spherical_probes_match_oracle(engine, oracle):
    """Phase 1 with spherical probes (FP64 probe operand, per-block column rescale)."""
    _say("phase1 spherical")
    m, n, K, V, step = 450, 290, 2, 5, 100
    codes = random_codes(m, n, seed=13, miss_rate=0.02)
    rng = np.random.default_rng(6)
    annot = np.ones((m, K))
    annot[:, 1] = rng.random(m) < 0.4
    inv = rng.uniform(0.8, 1.2, size=m)
    g = _geno(engine, oracle, codes).prepare(impute="mean")
    xz = torch.zeros((K * V, n), dtype=torch.float64, device="cuda")
    engine.phase1_xz(g, 0, m, step, torch.from_numpy(annot).cuda(), torch.from_numpy(inv).cuda(), 3, V, 4242, xz,
                     rand_dist="spherical")
    bed = oracle.pack_codes(codes)
    ref = np.zeros((n, K * V))
    for s in range(0, m, step):
        e = min(m, s + step)
        X = oracle.read_block_standardized(bed, n, s, e, None, 1, "mean", 0)
        ref += oracle.phase1_chunk(X, annot[s:e], inv[s:e], oracle.probe_block(4242, s, 3, e - s, V, "spherical"))
    # real-valued probes: the 38-bit column quantization is relative to the largest
    # |w z| of the column, which Gaussian tails push to several times the typical one
    err = np.abs(xz.cpu().numpy().T - ref) / np.abs(ref).max(axis=0, keepdims=True)
    assert err.max() < 3e-11, err.max()
    g.close()


def _phase1_ref(oracle, codes, annot, inv, seed, v_start, v, step, mode="mean", impute_seed=0):
    bed = oracle.pack_codes(codes)
    m, n = codes.shape
    xz = np.zeros((n, annot.shape[1] * v))
    for s in range(0, m, step):
        e = min(m, s + step)
        X = oracle.read_block_standardized(bed, n, s, e, None, 1, mode, impute_seed)
        Z = oracle.probe_block(seed, s, v_start, e - s, v)
        xz += oracle.phase1_chunk(X, annot[s:e], inv[s:e], Z)
    return xz


@pytest.mark.parametrize("mode", ["mean", "hwe"])
def test_phase1_matches_oracle(engine, oracle, mode, K=3, V=7, method=0):
    """method: missing-call term from per-call records (1), the missing-flag
    plane (2), or the built-in rule (0)."""
    from summit_amd import autotune
    _say(f"phase1 {mode} bins={K} probes={V} method={method}")
    m, n, step = 700, 333, 100
    autotune.set_value("p1_miss", autotune.shape_key(n, m, K, V, 5), method)
    codes = random_codes(m, n, seed=11, miss_rate=0.03, monomorphic=2)
    rng = np.random.default_rng(3)
    annot = (rng.random((m, K)) < 0.6).astype(np.float64) * rng.uniform(0.5, 2.0, size=(m, K))
    annot[:, 0] = 1.0
    inv = rng.uniform(0.8, 1.2, size=m)
    inv[5] = 0.0  # skipped by the bin scale rule
    iseed = 99
    g = _geno(engine, oracle, codes).prepare(impute=mode, impute_seed=iseed)
    xz = torch.zeros((K * V, n), dtype=torch.float64, device="cuda")
    engine.phase1_xz(g, 0, m, step, torch.from_numpy(annot).cuda(), torch.from_numpy(inv).cuda(), 5, V, 777, xz)
    got = xz.cpu().numpy().T
    ref = _phase1_ref(oracle, codes, annot, inv, 777, 5, V, step, mode, iseed)
    scale = np.abs(ref).max(axis=0, keepdims=True)
    err = np.abs(got - ref) / scale
    autotune.set_value("p1_miss", autotune.shape_key(n, m, K, V, 5), 0)
    assert err.max() < 1e-11, err.max()
    g.close()


@pytest.mark.parametrize("method", [1, 2])
def test_phase1_missing_methods_match_oracle(engine, oracle, method):
    """Per-call records and the missing-flag plane, each against the oracle,
    with several bins and 600 probes (ten 64-column words, padded columns)."""
    test_phase1_matches_oracle(engine, oracle, "mean", 3, 7, method)
    test_phase1_matches_oracle(engine, oracle, "mean", 2, 600, method)


@pytest.mark.parametrize("impute", ["mean", "hwe"])
def test_tiled_and_row_major_layouts_agree(engine, oracle, monkeypatch, impute):
    """The contractions read chunk-tiled copies of both packed layouts when they
    fit in device memory and the row-major layouts otherwise (SUMMIT_AMD_TILED=0
    forces that): phase 1, phase 2 and the exact integer products are
    bit-identical either way, with a ragged last row block (N = 1100, M = 2900)."""
    m, n, V = 2900, 1100, 40
    _say(f"tiled vs row-major layouts N={n} M={m} {impute}")
    codes = random_codes(m, n, seed=41, miss_rate=0.01)
    outs = []
    for flag in ("0", "1"):
        monkeypatch.setenv("SUMMIT_AMD_TILED", flag)
        g = _geno(engine, oracle, codes).prepare(impute=impute, impute_seed=5)
        annot = torch.ones((m, 1), dtype=torch.float64, device="cuda")
        inv = engine.resvar_inv(g, 0, m)
        xz = torch.zeros((V, n), dtype=torch.float64, device="cuda")
        engine.phase1_xz(g, 0, m, 500, annot, inv, 0, V, 17, xz)
        ms = torch.zeros((m, 1), dtype=torch.float64, device="cuda")
        engine.phase2_accum(g, 0, m, inv, V, xz, ms, n)
        z = (torch.randint(0, 2, (m, 64), generator=torch.Generator().manual_seed(3)) * 2 - 1).to(torch.int8).cuda()
        u = engine.gz_i32(g, z)
        outs.append((xz.cpu().numpy(), ms.cpu().numpy(), u.cpu().numpy()))
        g.close()
    monkeypatch.delenv("SUMMIT_AMD_TILED")
    for a, b in zip(outs[0], outs[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("K", [7, 17])
def test_phase1_many_bins_match_oracle(engine, oracle, K):
    """Bin counts the missing-call finalize splits into groups (7 = 4 + 2 + 1, 17 = 16 + 1)."""
    test_phase1_matches_oracle(engine, oracle, "mean", K, method=1)


def test_phase1_wide_probe_tile_matches_oracle(engine, oracle):
    """100 probes: two 64-column sign words per missing-call record, padded columns."""
    test_phase1_matches_oracle(engine, oracle, "mean", 2, 100, method=1)


@pytest.mark.parametrize("mode,limbs,rtol", [("mean", 5, 5e-11), ("hwe", 5, 5e-11), ("mean", 6, 1e-12)])
def test_phase2_matches_oracle(engine, oracle, mode, limbs, rtol):
    """Tolerance: 5 limbs quantize X.Z to 38 bits per column, 6 limbs to 44."""
    _say(f"phase2 {mode} limbs={limbs}")
    m, n, K, V = 500, 411, 2, 9
    codes = random_codes(m, n, seed=21, miss_rate=0.03, monomorphic=2, all_missing=1)
    rng = np.random.default_rng(4)
    xz = rng.standard_normal((n, K * V)) * 30.0
    inv = rng.uniform(0.8, 1.2, size=m)
    g = _geno(engine, oracle, codes).prepare(impute=mode, impute_seed=5)
    meansq = torch.zeros((m, K), dtype=torch.float64, device="cuda")
    mc_agg = torch.zeros(K, dtype=torch.float64, device="cuda")
    mc_ps = torch.zeros((m, K), dtype=torch.float64, device="cuda")
    lo, hi = 100, 431
    engine.phase2_accum(g, lo, hi, torch.from_numpy(inv[lo:hi]).cuda(), V,
                        torch.from_numpy(np.ascontiguousarray(xz.T)).cuda(), meansq, n_denom=n,
                        mc_aggregate_sumsq=mc_agg, mc_per_snp_sumsq=mc_ps, limbs=limbs)
    bed = oracle.pack_codes(codes)
    X = oracle.read_block_standardized(bed, n, lo, hi, None, 1, mode, 5)
    ref = np.zeros((m, K))
    ref_agg = np.zeros(K)
    ref_ps = np.zeros((m, K))
    oracle.phase2_accum(X, inv[lo:hi], V, np.asfortranarray(xz), ref, n, lo, ref_agg, ref_ps)
    got = meansq.cpu().numpy()
    assert np.all(got[:lo] == 0) and np.all(got[hi:] == 0)
    np.testing.assert_allclose(got, ref, rtol=rtol, atol=1e-300)
    np.testing.assert_allclose(mc_ps.cpu().numpy(), ref_ps, rtol=4 * rtol, atol=1e-300)
    np.testing.assert_allclose(mc_agg.cpu().numpy(), ref_agg, rtol=4 * rtol)
    g.close()


@pytest.mark.parametrize("mode,limbs", [("mean", 5), ("hwe", 5), ("mean", 6)])
def test_ldscore_matches_oracle(engine, oracle, mode, limbs):
    _say(f"ldscore {mode} limbs={limbs}")
    from summit_amd import ldscore
    m, n, nvecs, step = 1500, 600, 20, 500
    codes = random_codes(m, n, seed=31, miss_rate=0.01)
    seed = 1101
    iseed = ldscore.impute_seed_from_root(seed)
    g = _geno(engine, oracle, codes).prepare(impute=mode, impute_seed=iseed)
    res = ldscore.genomewide_ldscore(g, nvecs, seed, step_size=step, vtiles=[12, 8], limbs=limbs)
    ref = oracle.gw_ldscore(oracle.pack_codes(codes), n, m, nvecs=nvecs, step_size=step, seed=seed,
                            impute=mode, vtiles=[12, 8])
    np.testing.assert_allclose(res.inv, ref.inv, rtol=1e-13)
    np.testing.assert_allclose(res.meansq_accum, ref.meansq_accum, rtol=1e-11)
    assert_ld_close(res.ldscore, ref.ldscore, np.full(1, m / (n - 1.0)))
    g.close()


@pytest.mark.parametrize("rand_dist", ["rademacher", "gaussian"])
def test_wide_traversals_keep_the_vtile_streams(engine, oracle, rand_dist):
    """Several equal V-tiles served by one traversal pair (probe columns in tiles,
    each tile with its own per-block streams) give the reference's estimate for
    that V-tile split: tiles [8, 8, 8, 4] in traversals of 24 + 4 columns; a
    SNP block count (1500 / 400 -> 4 blocks, the last short) with several
    streams per launch; bins."""
    _say(f"wide traversals {rand_dist}")
    from summit_amd import ldscore
    m, n, step = 1500, 600, 400
    codes = random_codes(m, n, seed=33, miss_rate=0.01)
    seed = 77
    annot = np.ones((m, 2))
    annot[:, 1] = np.random.default_rng(4).random(m) < 0.4
    g = _geno(engine, oracle, codes).prepare(impute="mean")
    vt = [8, 8, 8, 4]
    assert ldscore.traversal_plan(vt, 24) == [(0, 24, 8), (24, 4, 4)]
    res = ldscore.genomewide_ldscore(g, 28, seed, annot=annot, step_size=step, vtiles=vt, traversal_width=24,
                                     rand_dist=rand_dist)
    ref = oracle.gw_ldscore(oracle.pack_codes(codes), n, m, annot=annot, nvecs=28, step_size=step, seed=seed,
                            impute="mean", vtiles=vt, rand_dist=rand_dist)
    assert_ld_close(res.ldscore, ref.ldscore, annot.sum(axis=0) / (n - 1.0))
    one = ldscore.genomewide_ldscore(g, 28, seed, annot=annot, step_size=step, vtiles=vt, rand_dist=rand_dist)
    np.testing.assert_allclose(res.meansq_accum, one.meansq_accum, rtol=1e-12)
    g.close()
