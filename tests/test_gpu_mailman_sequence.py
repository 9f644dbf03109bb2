"""The drop-in driven through the reference's Mailman backend sequence
(src/summit/ldscore/gw_ldscore.py:1934-1958 phase 1 row-major, :2010-2016
column sums / in-place projection, :2063-2079 Mailman phase 2 with sum_Xz),
which the reference takes with HWE imputation and at most 10 probes
(:665-684); checked against the oracle."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import assert_ld_close, random_codes

pytestmark = pytest.mark.gpu


def _write(oracle, codes, name):
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, name)
    oracle.write_bed(prefix, codes)
    return prefix


def _mailman_driver(gw, prefix, n, m, annot, nvecs, step, seed, C=None, R=None, sums_shift=None):
    """One V-tile of the reference's Mailman loop: phase 1 row-major per block with
    the fused residual variances (and their covariates), column sums (after the
    in-place projection when C is given), phase 2 Mailman row-major per block."""
    impute_seed = (seed ^ 0xA24BAED4963EE407) & 0xFFFFFFFFFFFFFFFF
    B = annot.shape[1]
    p = 0 if C is None else C.shape[1]
    n_denom = n - p if C is not None else n
    blocks = [(s, min(m, s + step)) for s in range(0, m, step)]
    inv_all = np.empty(m)
    meansq = np.zeros((m, B))
    mc_ps = np.zeros((m, B))
    mc_agg = np.zeros(B)
    xz = np.zeros((n, B * nvecs))  # C-contiguous
    for s, e in blocks:
        gw.phase1_compute_Xz_bed_chunk_rowmajor(
            bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=s, blk_end=e, row_sel=None, ddof=1,
            annot_blk=np.ascontiguousarray(annot[s:e]), inv_right=inv_all[s:e], v_start=0, v_count=nvecs,
            kmax_hint=e - s, rand_dist="rademacher", seed=seed, Xz2d_chunk=xz, project_right=False, C=C, R=R,
            impute_mode="hwe", impute_seed=impute_seed, resvar_gram=None, inv_out=inv_all[s:e], resvar_eps=1e-10)
    if C is not None:
        sums = gw.project_rowmajor_inplace_and_col_sums(xz, C, R)
    else:
        sums = gw.compute_col_sums_rowmajor(xz)
    if sums_shift is not None:
        sums = sums + sums_shift
    for s, e in blocks:
        gw.phase2_accum_XtXz_bed_mailman_rowmajor(
            bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=s, blk_end=e, row_sel=None, ddof=1,
            inv_left=inv_all[s:e], tile_nvecs=nvecs, Xz2d=xz, meansq_accum=meansq, mc_aggregate_sumsq=mc_agg,
            mc_per_snp_sumsq=mc_ps, sum_Xz=sums, N_denom=n_denom, impute_seed=impute_seed)
    null = annot.sum(axis=0) / float(n_denom - 1)
    ld = meansq / float(nvecs) - null[None, :]
    return ld, meansq, inv_all, xz, sums, mc_ps, mc_agg, null


def test_mailman_sequence_c1_shape(oracle):
    """N=5,000, M=10,000, B=10 Rademacher, i.i.d. HWE genotypes with 1% missing
    calls completed by the HWE hash, blocks of 1,000."""
    print("mailman sequence: N=5000 M=10000 B=10 hwe", flush=True)
    from summit_amd import gwldcore as gw
    n, m, nvecs, step, seed = 5000, 10000, 10, 1000, 1101
    codes = random_codes(m, n, seed=101, miss_rate=0.01)
    prefix = _write(oracle, codes, "c1")
    annot = np.ones((m, 1))
    ld, meansq, inv, xz, sums, mc_ps, mc_agg, null = _mailman_driver(gw, prefix, n, m, annot, nvecs, step, seed)
    ref = oracle.gw_ldscore(oracle.pack_codes(codes), n, m, nvecs=nvecs, step_size=step, seed=seed, impute="hwe",
                            mc=True)
    np.testing.assert_allclose(inv, ref.inv, rtol=1e-13)
    # the column sums the caller passes are those of X.Z
    np.testing.assert_allclose(sums, xz.sum(axis=0), rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(meansq, ref.meansq_accum, rtol=5e-11)
    np.testing.assert_allclose(mc_ps, ref.mc_per_snp_sumsq, rtol=2e-10)
    np.testing.assert_allclose(mc_agg, ref.mc_per_snp_sumsq.sum(axis=0), rtol=1e-10)
    assert_ld_close(ld, ref.ldscore, null)


def test_mailman_sequence_with_covariates(oracle):
    """Covariates: fused residual scales with C/R in phase 1, X.Z projected in
    place with its column sums, phase 2 without C/R on the projected X.Z."""
    from summit_amd import gwldcore as gw
    n, m, nvecs, step, seed = 600, 700, 8, 250, 77
    codes = random_codes(m, n, seed=102, miss_rate=0.02)
    prefix = _write(oracle, codes, "cov")
    rng = np.random.default_rng(6)
    Q, _ = np.linalg.qr(np.column_stack([np.ones(n), rng.standard_normal((n, 2))]))
    C = np.asfortranarray(Q)
    R = np.asfortranarray(Q.T)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.4).astype(np.float64)])
    ld, meansq, inv, xz, sums, mc_ps, _, null = _mailman_driver(gw, prefix, n, m, annot, nvecs, step, seed, C, R)
    # oracle, piece by piece: standardized HWE blocks, residual scales of P x,
    # phase 1 unprojected, (I - C R) X.Z, dense phase 2 (exact sums)
    impute_seed = (seed ^ 0xA24BAED4963EE407) & 0xFFFFFFFFFFFFFFFF
    bed = oracle.pack_codes(codes)
    X = oracle.read_block_standardized(bed, n, 0, m, None, 1, "hwe", impute_seed)
    Xp = X - C @ (R @ X)
    inv_ref = 1.0 / np.sqrt(np.maximum(np.einsum("ij,ij->j", Xp, Xp) / (n - 3 - 1), 1e-10))
    np.testing.assert_allclose(inv, inv_ref, rtol=1e-11)
    y = np.zeros((n, 2 * nvecs), order="F")
    for s in range(0, m, step):
        e = min(m, s + step)
        Z = oracle.probe_block(seed, s, 0, e - s, nvecs)
        y += oracle.phase1_chunk(X[:, s:e], annot[s:e], inv_ref[s:e], Z)
    y = y - C @ (R @ y)
    np.testing.assert_allclose(xz, y, rtol=1e-10, atol=1e-10 * np.max(np.abs(y)))
    ref = np.zeros((m, 2))
    mc_ref = np.zeros((m, 2))
    oracle.phase2_accum(np.asfortranarray(X), inv_ref, nvecs, np.asfortranarray(y), ref, n - 3, 0,
                        mc_per_snp_sumsq=mc_ref)
    np.testing.assert_allclose(meansq, ref, rtol=5e-11)
    np.testing.assert_allclose(mc_ps, mc_ref, rtol=2e-10)
    assert_ld_close(ld, ref / nvecs - null[None, :], null)


def test_mailman_phase2_centres_with_the_callers_sums(oracle):
    """sum_Xz enters the centring of every column (gwldcore.cpp:874-880): shifted
    sums give the oracle's answer for those sums, not the exact column sums'."""
    from summit_amd import gwldcore as gw
    n, m, nvecs, step, seed = 300, 400, 6, 200, 5
    codes = random_codes(m, n, seed=103, miss_rate=0.03)
    prefix = _write(oracle, codes, "sums")
    annot = np.ones((m, 1))
    shift = np.linspace(-3.0, 4.0, nvecs)
    ld, meansq, inv, xz, sums, mc_ps, _, _ = _mailman_driver(gw, prefix, n, m, annot, nvecs, step, seed,
                                                             sums_shift=shift)
    impute_seed = (seed ^ 0xA24BAED4963EE407) & 0xFFFFFFFFFFFFFFFF
    bed = oracle.pack_codes(codes)
    ref = np.zeros((m, 1))
    mc_ref = np.zeros((m, 1))
    for s in range(0, m, step):
        e = min(m, s + step)
        c, nobs, su, ss = oracle.decode_block(bed, n, s, e)
        _, mu, inv_std = oracle.standardize_block(c, nobs, su, ss, 1, "hwe", impute_seed, s)
        comp, _, _ = oracle.hwe_impute(c, nobs, su, impute_seed, s)
        G = np.array([0, 0, 1, 2], dtype=np.int64)[comp].T
        oracle.phase2_accum_mailman(G, mu, inv_std, inv[s:e], nvecs, xz, sums, ref, n, s, mc_per_snp_sumsq=mc_ref)
    np.testing.assert_allclose(meansq, ref, rtol=5e-11)
    np.testing.assert_allclose(mc_ps, mc_ref, rtol=2e-10)
    exact = np.zeros((m, 1))
    for s in range(0, m, step):
        e = min(m, s + step)
        gw.phase2_accum_XtXz_bed_mailman_rowmajor(
            bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=s, blk_end=e, inv_left=inv[s:e],
            tile_nvecs=nvecs, Xz2d=xz, meansq_accum=exact, sum_Xz=sums - shift, N_denom=n, impute_seed=impute_seed)
    assert np.max(np.abs(exact - meansq) / meansq) > 1e-3
    with pytest.raises(RuntimeError, match="sum_Xz shape mismatch"):
        gw.phase2_accum_XtXz_bed_mailman_rowmajor(
            bed_prefix=prefix, fam_path=prefix + ".fam", blk_start=0, blk_end=step, inv_left=inv[:step],
            tile_nvecs=nvecs, Xz2d=xz, meansq_accum=exact, sum_Xz=sums[:-1], N_denom=n, impute_seed=impute_seed)
