"""The drop-in ``summit_amd.gwldcore`` module driven the way the reference's
GenomewideLDScore drives ``summit.gwldcore`` (src/summit/ldscore/gw_ldscore.py:1905-2133):
per SNP block, per V-tile, NumPy arrays in and out, keyword arguments."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import assert_ld_close, random_codes

pytestmark = pytest.mark.gpu


def _write(oracle, codes, name="g"):
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, name)
    oracle.write_bed(prefix, codes)
    return prefix


def _reference_style_run(gw, prefix, n, m, annot, nvecs, step, seed, impute, vtiles, dtype=np.float64, rows=None):
    """The reference driver's loop over blocks and V-tiles (fused residual variances in tile 0)."""
    impute_seed = (seed ^ 0xA24BAED4963EE407) & 0xFFFFFFFFFFFFFFFF
    B = annot.shape[1]
    blocks = [(s, min(m, s + step)) for s in range(0, m, step)]
    inv_all = np.empty(m, dtype=dtype)
    meansq = np.zeros((m, B), dtype=dtype)
    mc_ps = np.zeros((m, B))
    nsel = n if rows is None else len(rows)
    v_start = 0
    for vt_idx, Vt in enumerate(vtiles):
        xz = np.empty((nsel, B * Vt), dtype=dtype, order="F")
        init = np.zeros(B, dtype=np.uint8)
        for s, e in blocks:
            gw.phase1_compute_Xz_bed_chunk(
                bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=s, blk_end=e, row_sel=rows, ddof=1,
                annot_blk=np.ascontiguousarray(annot[s:e].astype(dtype)), inv_right=inv_all[s:e], v_start=v_start,
                v_count=Vt, kmax_hint=e - s, rand_dist="rademacher", seed=seed, Xz2d_chunk=xz, bin_init_mask=init,
                project_right=False, impute_mode=impute, impute_seed=impute_seed,
                inv_out=inv_all[s:e] if vt_idx == 0 else None, resvar_eps=1e-10)
        for s, e in blocks:
            gw.phase2_accum_XtXz_bed(
                bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=s, blk_end=e, row_sel=rows, ddof=1,
                inv_left=inv_all[s:e], tile_nvecs=Vt, Xz2d=xz, meansq_accum=meansq, N_denom=nsel,
                impute_mode=impute, impute_seed=impute_seed, mc_per_snp_sumsq=mc_ps)
        v_start += Vt
    ld = meansq / float(nvecs) - (annot.sum(axis=0) / float(nsel - 1))[None, :]
    return ld, meansq, inv_all, mc_ps


@pytest.mark.parametrize("impute", ["mean", "hwe"])
def test_dropin_driver_matches_oracle(oracle, impute):
    from summit_amd import gwldcore as gw
    m, n, step, nvecs, seed = 900, 301, 250, 12, 4242
    codes = random_codes(m, n, seed=77, miss_rate=0.02, monomorphic=2, all_missing=1)
    prefix = _write(oracle, codes)
    rng = np.random.default_rng(1)
    annot = np.ones((m, 2))
    annot[:, 1] = (rng.random(m) < 0.3)
    ld, meansq, inv, mc = _reference_style_run(gw, prefix, n, m, annot, nvecs, step, seed, impute, [8, 4])
    ref = oracle.gw_ldscore(oracle.pack_codes(codes), n, m, annot=annot, nvecs=nvecs, step_size=step, seed=seed,
                            impute=impute, vtiles=[8, 4], mc=True)
    np.testing.assert_allclose(inv, ref.inv, rtol=1e-13)
    np.testing.assert_allclose(meansq, ref.meansq_accum, rtol=5e-11)
    np.testing.assert_allclose(mc, ref.mc_per_snp_sumsq, rtol=2e-10)
    assert_ld_close(ld, ref.ldscore, annot.sum(axis=0) / (n - 1.0))


def test_dropin_row_selection_and_float32(oracle):
    from summit_amd import gwldcore as gw
    m, n, step, nvecs, seed = 400, 257, 128, 8, 9
    codes = random_codes(m, n, seed=5, miss_rate=0.02)
    prefix = _write(oracle, codes, "rs")
    rows = np.sort(np.random.default_rng(3).choice(n, 190, replace=False)).astype(np.int64)
    annot = np.ones((m, 1))
    ld, meansq, inv, _ = _reference_style_run(gw, prefix, n, m, annot, nvecs, step, seed, "mean", [8], rows=rows)
    ref = oracle.gw_ldscore(oracle.pack_codes(codes), n, m, nvecs=nvecs, step_size=step, seed=seed, impute="mean",
                            rows=rows)
    assert_ld_close(ld, ref.ldscore, np.full(1, m / (190 - 1.0)))
    ld32, *_ = _reference_style_run(gw, prefix, n, m, annot, nvecs, step, seed, "mean", [8], dtype=np.float32,
                                    rows=rows)
    assert_ld_close(ld32, ref.ldscore, np.full(1, m / (190 - 1.0)), rtol=1e-6)


def test_dropin_errors_mirror_reference(oracle):
    from summit_amd import gwldcore as gw
    codes = random_codes(20, 16, seed=1)
    prefix = _write(oracle, codes, "err")
    xz = np.zeros((16, 4), order="F")
    kw = dict(bed_prefix=prefix, fam_path=prefix + ".fam", row_sel=None, ddof=1, annot_blk=np.ones((10, 1)),
              inv_right=np.ones(10), v_start=0, v_count=4, kmax_hint=10, seed=1, Xz2d_chunk=xz, impute_mode="mean")
    with pytest.raises(RuntimeError, match="blk_end exceeds"):
        gw.phase1_compute_Xz_bed_chunk(blk_start=15, blk_end=25, **kw)
    with pytest.raises(RuntimeError, match="impute_mode must be"):
        gw.phase1_compute_Xz_bed_chunk(blk_start=0, blk_end=10, **{**kw, "impute_mode": "zero"})
    with pytest.raises(RuntimeError, match="row_sel contains an out-of-range"):
        gw.phase1_compute_Xz_bed_chunk(blk_start=0, blk_end=10, **{**kw, "row_sel": np.array([0, 99])})
    with pytest.raises(RuntimeError, match="L mismatch"):
        gw.phase1_compute_Xz_bed_chunk(blk_start=0, blk_end=9, **kw)
    with pytest.raises(RuntimeError, match="multiple of tile_nvecs"):
        gw.phase2_accum_XtXz_bed(bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=0, blk_end=10,
                                 inv_left=np.ones(10), tile_nvecs=3, Xz2d=xz, meansq_accum=np.zeros((20, 1)),
                                 N_denom=16, impute_mode="mean")
    with open(prefix + ".bed", "r+b") as f:
        f.write(b"\x00\x00\x00")
    gw._clear_bed_mapping_cache()
    with pytest.raises(RuntimeError, match="Invalid BED magic"):
        gw.phase1_compute_Xz_bed_chunk(blk_start=0, blk_end=10, **kw)


def test_covariate_projection_matches_dense_oracle(oracle):
    """Phase 2 with C/R projects the genotype block (gwldcore.cpp:2374-2388); residual scales
    with covariates use ||P x||^2 / (N - p - 1) (gwldcore.cpp:3262-3332)."""
    from summit_amd import gwldcore as gw
    m, n, V = 120, 200, 5
    codes = random_codes(m, n, seed=8, miss_rate=0.02)
    prefix = _write(oracle, codes, "cov")
    rng = np.random.default_rng(2)
    # the reference passes an orthonormal covariate basis C = Q and R = Q^T (gw_ldscore.py:1015-1035)
    Q, _ = np.linalg.qr(np.column_stack([np.ones(n), rng.standard_normal((n, 2))]))
    C = np.asfortranarray(Q)
    R = np.asfortranarray(Q.T)
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, "mean", 0)
    Xp = X - C @ (R @ X)
    xz = np.asfortranarray(rng.standard_normal((n, V)))
    inv = rng.uniform(0.9, 1.1, size=m)
    meansq = np.zeros((m, 1))
    gw.phase2_accum_XtXz_bed(bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=0, blk_end=m, inv_left=inv,
                             tile_nvecs=V, Xz2d=xz, meansq_accum=meansq, C=C, R=R, N_denom=n - 3,
                             impute_mode="mean")
    ref = np.zeros((m, 1))
    oracle.phase2_accum(np.asfortranarray(Xp), inv, V, xz, ref, n - 3, 0)
    np.testing.assert_allclose(meansq, ref, rtol=1e-10)
    invc, _ = gw.precompute_residual_variances_bed(prefix, prefix + ".fam", m, 50, None, 1, 1e-10, False, None, C, R,
                                                   "mean", None)
    ss = np.einsum("ij,ij->j", Xp, Xp)
    np.testing.assert_allclose(invc, 1.0 / np.sqrt(np.maximum(ss / (n - 3 - 1), 1e-10)), rtol=1e-11)
