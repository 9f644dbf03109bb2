"""ORACLE — CPU restatement of SUMMIT's genome-wide stochastic LD-score engine.

This module is test infrastructure. Only ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` may import it, and only as the
checker or the timed CPU baseline; the product (``summit_amd``) never does.

Every function restates one piece of the reference and cites it. Paths are
relative to the reference repository root.

Pins (each function is checked by something other than a retyping of itself):

* BED layout / decode / stats / mean and HWE standardization
  (``decode_block``, ``hwe_impute``, ``standardize_block``): checked against the
  reference's own ``read_block_standardized_double`` compiled from
  ``src/native/common/genotype.cpp`` (``oracle/_ref``), live where the
  reference tree exists and through committed golden vectors elsewhere, and
  against the textbook column standardization of
  ``tests/test_pgen_gwld.py:244-262``.
* Probe streams (``probe_block``): MT19937-64 known answer from the C++
  standard; libstdc++'s own ``std::mt19937_64``/``std::normal_distribution``
  through ``oracle/_ref``; golden vectors.
* Phase 1 / phase 2 / finalize (``phase1_chunk``, ``phase2_accum``,
  ``gw_ldscore``): the dense matrix oracle of ``tests/test_gwld_mc.py:42-128``
  (hand-checked small matrices), invariance of the estimate to block size and
  V-tiling where the reference says it is invariant, and the statistical pin of
  ``tests/test_pgen_gwld.py:715-772`` (error against exact LD within 5 Monte
  Carlo standard errors, variance ``2[(sum r^2)^2 - sum r^4]/B``).
* Integer products (``gz_int``, ``gtu_int``): brute-force int64 loops on tiny
  inputs and the identity ``sum_b U[:,b]^2 == (Z^T G G^T Z)`` trace checks.
* GxE features (``gxe_project_center`` and the ``inv_x``/``inv_w`` of
  ``gxe_scores``): the reference test's dense feature formula
  (``tests/test_gxe_native_core.py:95-123``: ``p = I - q q^T`` with the
  intercept in ``q``, rank ``N - q.shape[1]``), the invariants ``C^T x = 0``,
  ``1^T x = 0``, ``||x_j||^2 / (N - p - 1) = 1``, and planted mistakes (no
  centring, environment applied after the projection, a dropped degree of
  freedom) that the formula rejects.
* GxE DirectContext (``direct_decode``, ``direct_features``, ``direct_source``,
  ``direct_target``, ``direct_leakage``): the decode against the reference
  test's mean-filled ``std(ddof)`` route (``tests/test_gxe_native_core.py:
  95-123``) and against the genome-wide decode up to sign and the
  ``sqrt((N - ddof) / (nobs - ddof))`` factor; the dense features against the
  reference's moment algebra (``Q^T E^p X`` moments, gxeldcore.cpp:8033-8243)
  restated independently in the test; invariants (``norm_x = norm_w = 1``,
  ``q^T source = 0``, ``target_x = diag(sx) X^T P X diag(a sx) Z`` through the
  dense Gram); brute-force loops on a tiny block; planted mistakes (environment
  applied after the projection, ``nobs - ddof`` in place of ``N - ddof``,
  unprojected targets) that the checks reject.
* Fused residual-variance pass (``resvar_pass``): hand-computed mu22 of one SNP
  with and without a missing call, invariance to SNP duplication and block
  size, the Cauchy-Schwarz bound of complete data, the (N - 2)/(N - 1) scaling
  of an intercept-only projection, and V against the reference test's formula
  (``tests/test_snp_var_weights.py:52-67``).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

U64 = np.uint64
MASK53 = U64((1 << 53) - 1)


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libgpe_oracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle C library missing: build it with `make -C {_HERE}`")
        lib = ctypes.CDLL(path)
        lib.or_mt64_nth.restype = ctypes.c_uint64
        lib.or_mt64_nth.argtypes = [ctypes.c_uint64, ctypes.c_int64]
        lib.or_make_seed.restype = ctypes.c_uint64
        lib.or_make_seed.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int]
        lib.or_mailman_segment_size.restype = ctypes.c_int
        lib.or_mailman_segment_size.argtypes = [ctypes.c_int64]
        for fn in (lib.or_mailman_gz, lib.or_mailman_gtu):
            fn.restype = None
            fn.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                           ctypes.c_void_p, ctypes.c_int]
        for fn in (lib.or_rademacher_block, lib.or_gaussian_block):
            fn.restype = None
            fn.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                           ctypes.c_void_p]
        lib.or_real_probe_block.restype = None
        lib.or_real_probe_block.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_void_p]
        lib.or_stream_raw.restype = None
        lib.or_stream_raw.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int64,
                                      ctypes.c_void_p]
        _LIB = lib
    return _LIB


# ----------------------------------------------------------------------------
# PLINK BED (SNP-major, magic 6C 1B 01, 4 samples per byte, low bits first).
# Codes: 00 -> 0, 01 -> missing, 10 -> 1, 11 -> 2 copies of the counted allele
# (src/native/common/genotype.cpp:135-144, :384-389).
# ----------------------------------------------------------------------------
BED_MAGIC = bytes([0x6C, 0x1B, 0x01])
CODE_MISSING = 1


def geno_to_codes(g: np.ndarray) -> np.ndarray:
    """Map values {0,1,2,-9/negative=missing} to 2-bit codes (genotype.cpp:142-144)."""
    g = np.asarray(g)
    codes = np.full(g.shape, CODE_MISSING, dtype=np.uint8)
    codes[g == 0] = 0
    codes[g == 1] = 2
    codes[g == 2] = 3
    return codes


def codes_to_geno(codes: np.ndarray) -> np.ndarray:
    """2-bit codes to int8 genotype values with missing -> -9."""
    lut = np.array([0, -9, 1, 2], dtype=np.int8)
    return lut[codes]


def pack_codes(codes_snp_major: np.ndarray) -> np.ndarray:
    """codes (M, N) uint8 -> packed (M, ceil(N/4)) uint8."""
    c = np.asarray(codes_snp_major, dtype=np.uint8)
    M, N = c.shape
    bps = (N + 3) // 4
    pad = np.zeros((M, bps * 4), dtype=np.uint8)
    pad[:, :N] = c
    q = pad.reshape(M, bps, 4)
    return (q[:, :, 0] | (q[:, :, 1] << 2) | (q[:, :, 2] << 4) | (q[:, :, 3] << 6)).astype(np.uint8)


def unpack_codes(packed: np.ndarray, N: int) -> np.ndarray:
    """packed (M, bps) -> codes (M, N)."""
    p = np.asarray(packed, dtype=np.uint8)
    out = np.empty((p.shape[0], p.shape[1] * 4), dtype=np.uint8)
    for t in range(4):
        out[:, t::4] = (p >> (2 * t)) & 3
    return out[:, :N]


def write_bed(prefix: str, codes_snp_major: np.ndarray, chrom=None, bp=None) -> None:
    """Write .bed/.bim/.fam for codes (M, N)."""
    M, N = codes_snp_major.shape
    with open(prefix + ".bed", "wb") as f:
        f.write(BED_MAGIC)
        f.write(pack_codes(codes_snp_major).tobytes())
    if chrom is None:
        chrom = np.ones(M, dtype=np.int64)
    if bp is None:
        bp = np.arange(1, M + 1, dtype=np.int64) * 1000
    with open(prefix + ".bim", "w") as f:
        for j in range(M):
            f.write(f"{int(chrom[j])}\trs{j}\t0\t{int(bp[j])}\tA\tG\n")
    with open(prefix + ".fam", "w") as f:
        for n in range(N):
            f.write(f"f{n}\ti{n}\t0\t0\t0\t-9\n")


def read_bed_matrix(prefix: str, n_total: int, m_total: int) -> np.ndarray:
    raw = np.fromfile(prefix + ".bed", dtype=np.uint8)
    if raw.size < 3 or bytes(raw[:3]) != BED_MAGIC:
        raise RuntimeError("Invalid BED magic/mode (expected 6C 1B 01)")
    bps = (n_total + 3) // 4
    return raw[3:3 + bps * m_total].reshape(m_total, bps)


# ----------------------------------------------------------------------------
# Decode + stats: src/native/common/genotype.cpp:474-574
# ----------------------------------------------------------------------------
def decode_block(bed: np.ndarray, n_total: int, blk_start: int, blk_end: int, rows=None):
    """Returns codes (L, N_sel) and per-SNP (nobs, sum, sumsq) as int64."""
    codes = unpack_codes(bed[blk_start:blk_end], n_total)
    if rows is not None:
        codes = codes[:, np.asarray(rows, dtype=np.int64)]
    obs = codes != CODE_MISSING
    val = np.array([0, 0, 1, 2], dtype=np.int64)[codes]
    nobs = obs.sum(axis=1).astype(np.int64)
    s = val.sum(axis=1).astype(np.int64)
    ss = (val * val).sum(axis=1).astype(np.int64)
    return codes, nobs, s, ss


# ----------------------------------------------------------------------------
# HWE hash imputation: src/native/common/genotype.cpp:122-184, :576-595
# ----------------------------------------------------------------------------
def mix64(x):
    x = np.asarray(x, dtype=U64)
    with np.errstate(over="ignore"):
        x = x + U64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> U64(30))) * U64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> U64(27))) * U64(0x94D049BB133111EB)
    return x ^ (x >> U64(31))


def u01_from_u64(x):
    return ((np.asarray(x, dtype=U64) >> U64(11)) & MASK53).astype(np.float64) * (1.0 / 9007199254740992.0)


def hwe_seed_base(impute_seed: int, global_snp):
    with np.errstate(over="ignore"):
        h = mix64(U64(impute_seed))
        gs = np.asarray(global_snp, dtype=np.int64).astype(np.uint32).astype(U64)
        return h ^ mix64(gs + U64(0x243F6A8885A308D3))


def hwe_thresholds(obs_sum, nobs):
    obs_sum = np.asarray(obs_sum, dtype=np.float64)
    nobs = np.asarray(nobs, dtype=np.float64)
    p = np.where(nobs > 0, obs_sum / np.maximum(2.0 * nobs, 1.0), 0.0)
    p = np.clip(p, 0.0, 1.0)
    q0 = (1.0 - p) * (1.0 - p)
    q01 = q0 + 2.0 * p * (1.0 - p)
    return q0, q01


def hwe_impute(codes, nobs, s, impute_seed: int, blk_start: int, rows=None):
    """Complete missing codes by the deterministic HWE hash; returns completed
    codes and the completed (sum, sumsq)."""
    codes = codes.copy()
    L, N = codes.shape
    row_ids = np.arange(N, dtype=np.int64) if rows is None else np.asarray(rows, dtype=np.int64)
    q0, q01 = hwe_thresholds(s, nobs)
    for j in range(L):
        miss = np.flatnonzero(codes[j] == CODE_MISSING)
        if miss.size == 0:
            continue
        if q0[j] >= 1.0:
            vals = np.zeros(miss.size, dtype=np.int64)
        elif q01[j] <= 0.0:
            vals = np.full(miss.size, 2, dtype=np.int64)
        else:
            base = hwe_seed_base(impute_seed, blk_start + j)
            with np.errstate(over="ignore"):
                gr = row_ids[miss].astype(np.uint32).astype(U64)
                h = base ^ mix64(gr + U64(0x9E3779B97F4A7C15))
            u = u01_from_u64(mix64(h))
            vals = np.where(u < q0[j], 0, np.where(u < q01[j], 1, 2)).astype(np.int64)
        codes[j, miss] = np.array([0, 2, 3], dtype=np.uint8)[vals]
    val = np.array([0, 0, 1, 2], dtype=np.int64)[codes]
    return codes, val.sum(axis=1), (val * val).sum(axis=1)


# ----------------------------------------------------------------------------
# Standardization: src/native/common/genotype.cpp:629-689, :802-838
# ----------------------------------------------------------------------------
def _mean_invsd(nobs, s, ss, ddof):
    nobs = np.asarray(nobs, dtype=np.int64)
    dn = nobs.astype(np.float64)
    sf = np.asarray(s, dtype=np.float64)
    ssf = np.asarray(ss, dtype=np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        mu = np.where(nobs > 0, sf / np.where(nobs > 0, dn, 1.0), 0.0)
        M2 = np.where(nobs > 0, ssf - (sf * sf) / np.where(nobs > 0, dn, 1.0), 0.0)
    M2 = np.where((M2 < 0.0) & (M2 > -1e-12), 0.0, M2)
    denom = nobs - ddof
    inv = np.ones_like(mu)
    ok = (denom > 0) & (M2 > 0.0)
    var = np.where(ok, M2 / np.where(denom > 0, denom, 1).astype(np.float64), 0.0)
    ok &= var > 0.0
    inv[ok] = 1.0 / np.sqrt(var[ok])
    return mu, inv


def standardize_block(codes, nobs, s, ss, ddof: int, mode: str = "mean", impute_seed: int = 0,
                      blk_start: int = 0, rows=None):
    """Standardized block X (N, L) in Fortran order, plus per-SNP mean/inv_sd.

    mean mode: mu, inv_sd from observed calls; missing -> 0
    (genotype.cpp:652-689). hwe mode: complete by the HWE hash first, then mu,
    inv_sd over all N (genotype.cpp:802-830, :629-650)."""
    L, N = codes.shape
    if mode == "mean":
        mu, inv = _mean_invsd(nobs, s, ss, ddof)
        v = np.stack([(0.0 - mu) * inv, np.zeros_like(mu), (1.0 - mu) * inv, (2.0 - mu) * inv], axis=1)
        X = np.take_along_axis(v, codes.astype(np.int64), axis=1)
        return np.asfortranarray(X.T), mu, inv
    if mode == "hwe":
        comp, s2, ss2 = hwe_impute(codes, nobs, s, impute_seed, blk_start, rows)
        full = np.full(L, N, dtype=np.int64)
        mu, inv = _mean_invsd(full, s2, ss2, ddof)
        # N<=0 or N-ddof<=0 keeps inv 1 (genotype.cpp:636-638)
        v = np.stack([(0.0 - mu) * inv, np.zeros_like(mu), (1.0 - mu) * inv, (2.0 - mu) * inv], axis=1)
        X = np.take_along_axis(v, comp.astype(np.int64), axis=1)
        return np.asfortranarray(X.T), mu, inv
    raise ValueError("impute_mode must be 'mean' or 'hwe'")


def read_block_standardized(bed, n_total, blk_start, blk_end, rows=None, ddof=1, mode="mean",
                            impute_seed=0):
    codes, nobs, s, ss = decode_block(bed, n_total, blk_start, blk_end, rows)
    X, _, _ = standardize_block(codes, nobs, s, ss, ddof, mode, impute_seed, blk_start, rows)
    return X


# ----------------------------------------------------------------------------
# Probes: src/native/gwldcore.cpp:445-458, :1679-1716
# ----------------------------------------------------------------------------
def make_seed(root: int, block: int, v0: int) -> int:
    return int(_lib().or_make_seed(root, block, v0))


def mt64_nth(seed: int, n: int) -> int:
    return int(_lib().or_mt64_nth(seed, n))


def stream_raw(root: int, blk_start: int, v_start: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    _lib().or_stream_raw(root, blk_start, v_start, n, out.ctypes.data)
    return out


def probe_block(root: int, blk_start: int, v_start: int, L: int, V: int, dist: str = "rademacher",
                fp32: bool = False):
    """Z (L, V), column-major fill order, float64 values.  As the reference
    (src/native/gwldcore.cpp:1691-1716): "rademacher", "spherical", anything else
    Gaussian; fp32 draws the float-precision streams of float32 callers."""
    Z = np.empty((V, L), dtype=np.float64)
    if dist == "rademacher":
        _lib().or_rademacher_block(root, blk_start, v_start, L, V, Z.ctypes.data)
    else:
        code = (2 if dist == "spherical" else 1) + (2 if fp32 else 0)
        _lib().or_real_probe_block(root, blk_start, v_start, L, V, code, Z.ctypes.data)
    return Z.T  # (L, V), Fortran-contiguous


# ----------------------------------------------------------------------------
# Residual scale: src/native/gwldcore.cpp:3220-3333 (no covariates)
# ----------------------------------------------------------------------------
def inv_resvar(X, eps: float = 1e-10):
    N = X.shape[0]
    ss = np.einsum("ij,ij->j", X, X)
    var = ss / float(N - 1)
    safe = np.where(np.isnan(var), np.nan, np.maximum(var, eps))
    return 1.0 / np.sqrt(safe)


# ----------------------------------------------------------------------------
# Fused residual-variance pass: src/native/gwldcore.cpp:3335-3566
# ----------------------------------------------------------------------------
def resvar_pass(bed, n_total: int, nsnps: int, step_size: int, rows=None, ddof: int = 1, eps: float = 1e-10,
                annot=None, C=None, R=None, mode: str = "mean", impute_seed: int = 0):
    """(inv, mu22 or None, snp_var) of precompute_residual_variances_bed_impl:
    per block the standardized genotypes X, raw = ||x_j||^2, X <- X - C (R X)
    when C/R are given (:3468-3469), ss = ||x_j||^2 over proj_df = N - p - 1
    (N - 1 without covariates) into inv (eps floor, NaN kept, :3478-3490),
    snp_var = min(1, ss / raw) (0 where raw is 0 or ss is not finite, :3491-3493);
    with annot, S (N, B) += (X * X) (a * inv^2) (:3496-3521) and
    mu22[a, b] = (S^T S)[a, b] / (N Msum_a Msum_b), 0 where the denominator is 0
    or the value is not finite (:3528-3544)."""
    Nr = n_total if rows is None else len(rows)
    p = 0 if C is None else np.asarray(C).shape[1]
    proj_df = float(Nr - p - 1) if C is not None else float(Nr - 1)
    inv = np.empty(nsnps)
    snp_var = np.empty(nsnps)
    S = None
    if annot is not None:
        A = np.asarray(annot, dtype=np.float64)
        S = np.zeros((Nr, A.shape[1]))
    for s in range(0, nsnps, step_size):
        e = min(nsnps, s + step_size)
        X = np.array(read_block_standardized(bed, n_total, s, e, rows, ddof, mode, impute_seed), dtype=np.float64)
        raw = np.einsum("ij,ij->j", X, X)
        if C is not None:
            X = X - np.asarray(C, dtype=np.float64) @ (np.asarray(R, dtype=np.float64) @ X)
        ss = np.einsum("ij,ij->j", X, X)
        var = ss / proj_df
        safe = np.where(np.isnan(var), np.nan, np.maximum(var, eps))
        inv[s:e] = 1.0 / np.sqrt(safe)
        with np.errstate(divide="ignore", invalid="ignore"):
            snp_var[s:e] = np.where((raw > 0.0) & np.isfinite(ss), np.minimum(1.0, ss / raw), 0.0)
        if S is not None:
            S += (X * X) @ (A[s:e] * (inv[s:e] ** 2)[:, None])
    if S is None:
        return inv, None, snp_var
    msum = A[:nsnps].sum(axis=0)
    denom = float(Nr) * msum[:, None] * msum[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        mu22 = np.where(denom != 0.0, (S.T @ S) / denom, 0.0)
    mu22[~np.isfinite(mu22)] = 0.0
    return inv, mu22, snp_var


# ----------------------------------------------------------------------------
# Phase 1: src/native/gwldcore.cpp:1002-1063 (bin CSR scale), :1582-1861
# ----------------------------------------------------------------------------
def bin_scales(annot_blk, inv):
    """scale[i,k] = inv_i*sqrt(a_ik), 0 where the reference skips the entry."""
    a = np.asarray(annot_blk, dtype=np.float64)
    invd = np.asarray(inv, dtype=np.float64)[:, None]
    with np.errstate(invalid="ignore", over="ignore"):
        sc = invd * np.sqrt(np.where(a > 0, a, 0.0))
    ok = (a != 0) & (a > 0) & np.isfinite(a) & np.isfinite(invd) & (invd > 0) & np.isfinite(sc)
    return np.where(ok, sc, 0.0)


def phase1_chunk(X, annot_blk, inv, Z):
    """Xz contribution of one block: (N, B*V), column k*V+v."""
    sc = bin_scales(annot_blk, inv)  # (L, B)
    B = sc.shape[1]
    V = Z.shape[1]
    out = np.empty((X.shape[0], B * V), dtype=np.float64, order="F")
    for k in range(B):
        out[:, k * V:(k + 1) * V] = X @ (sc[:, k:k + 1] * Z)
    return out


# ----------------------------------------------------------------------------
# Phase 2: src/native/gwldcore.cpp:2329-2600 (+ guards :295-417)
# ----------------------------------------------------------------------------
EXTREME_LD_WARN = 1.0e10
QPANEL = 32768


def phase2_accum(X, inv_left, tile_nvecs: int, Xz, meansq_accum, N_denom: int, blk_start: int,
                 mc_aggregate_sumsq=None, mc_per_snp_sumsq=None, drop_extreme: bool = True):
    L = X.shape[1]
    Q = Xz.shape[1]
    if tile_nvecs <= 0 or Q % tile_nvecs:
        raise RuntimeError("Xz2d col count must be a multiple of tile_nvecs")
    B = Q // tile_nvecs
    denom = float(N_denom) - 1.0
    if denom <= 0.0:
        denom = 1.0
    invd2 = 1.0 / (denom * denom)
    inv = np.asarray(inv_left, dtype=np.float64)
    left = np.where(np.isfinite(inv) & (inv > 0), inv * inv * invd2, 0.0)
    qpanel = min(QPANEL, Q)
    qpanel = (qpanel // 64) * 64
    if qpanel < 64:
        qpanel = min(Q, 64)
    counts = {"nonfinite": 0, "extreme": 0}
    for q0 in range(0, Q, qpanel):
        q = min(qpanel, Q - q0)
        work = X.T @ Xz[:, q0:q0 + q]
        g = q0
        while g < q0 + q:
            k = g // tile_nvecs
            v_in = g - k * tile_nvecs
            ln = min(q0 + q - g, tile_nvecs - v_in)
            w = work[:, g - q0:g - q0 + ln]
            acc = np.sum(w * w, axis=1)
            upd = acc * left
            nonfin = ~np.isfinite(upd)
            extreme = np.isfinite(upd) & (upd > EXTREME_LD_WARN * float(ln))
            counts["nonfinite"] += int(nonfin.sum())
            counts["extreme"] += int(extreme.sum())
            keep = ~nonfin & ~(extreme & drop_extreme)
            upd = np.where(upd < 0.0, 0.0, upd)
            rows = blk_start + np.arange(L)
            meansq_accum[rows[keep], k] += upd[keep]
            if mc_aggregate_sumsq is not None or mc_per_snp_sumsq is not None:
                w2 = w * w
                sumsq = np.sum(w2 * w2, axis=1) * left * left
                if mc_per_snp_sumsq is not None:
                    mc_per_snp_sumsq[rows[keep], k] += sumsq[keep]
                if mc_aggregate_sumsq is not None:
                    mc_aggregate_sumsq[k] += float(np.sum(sumsq[keep]))
            g += ln
    return counts


# ----------------------------------------------------------------------------
# Driver: src/summit/ldscore/gw_ldscore.py:1753-2223 (BED, no covariates)
# ----------------------------------------------------------------------------
def phase2_accum_mailman(G, mean, inv_std, inv_left, tile_nvecs: int, Y, sum_y, meansq_accum, N_denom: int,
                         blk_start: int, mc_per_snp_sumsq=None, drop_extreme: bool = True):
    """Mailman phase 2 with the caller's column sums (src/native/gwldcore.cpp:2919-3119,
    per-SNP update :774-908): for HWE-completed counts G (N, L) and Y (N, Q),
    w = G^T Y (the exact products the 3^s tables give), and per SNP j and bin k
    upd = a_j sum_c w_jc^2 - 2 mu_j a_j sum_c w_jc s_c + mu_j^2 a_j sum_c s_c^2,
    a_j = inv_j^2 inv_std_j^2 / (N_denom - 1)^2, s = sum_y (the caller's sums),
    with the non-finite / extreme guards on the group length; the MC sum is
    a_j^2 sum_c (w_jc - mu_j s_c)^4."""
    G = np.asarray(G, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    s = np.asarray(sum_y, dtype=np.float64)
    L = G.shape[1]
    Q = Y.shape[1]
    if tile_nvecs <= 0 or Q % tile_nvecs:
        raise RuntimeError("Xz2d col count must be a multiple of tile_nvecs")
    if s.shape != (Q,):
        raise RuntimeError("sum_Xz shape mismatch in phase2_accum_XtXz_bed_mailman_rm")
    denom = float(N_denom) - 1.0
    if denom <= 0.0:
        denom = 1.0
    inv = np.asarray(inv_left, dtype=np.float64)
    alpha = np.where(np.isfinite(inv) & (inv > 0), inv * inv / (denom * denom), 0.0) * \
        np.asarray(inv_std, dtype=np.float64) ** 2
    mu = np.asarray(mean, dtype=np.float64)
    counts = {"nonfinite": 0, "extreme": 0}
    w = G.T @ Y  # (L, Q)
    B = Q // tile_nvecs
    for k in range(B):
        sl = slice(k * tile_nvecs, (k + 1) * tile_nvecs)
        wk, sk = w[:, sl], s[sl]
        upd = alpha * np.sum(wk * wk, axis=1) - 2.0 * mu * alpha * (wk @ sk) + mu * mu * alpha * float(sk @ sk)
        live = np.isfinite(alpha) & (alpha > 0)  # other SNPs are skipped before the guards
        nonfin = live & ~np.isfinite(upd)
        extreme = live & np.isfinite(upd) & (upd > EXTREME_LD_WARN * float(tile_nvecs))
        counts["nonfinite"] += int(nonfin.sum())
        counts["extreme"] += int(extreme.sum())
        keep = live & ~nonfin & ~(extreme & drop_extreme)
        upd = np.where(upd < 0.0, 0.0, upd)
        rows = blk_start + np.arange(L)
        meansq_accum[rows[keep], k] += upd[keep]
        if mc_per_snp_sumsq is not None:
            cen = wk - mu[:, None] * sk[None, :]
            c2 = cen * cen
            mc_per_snp_sumsq[rows[keep], k] += (alpha * alpha * np.sum(c2 * c2, axis=1))[keep]
    return counts


def bounded_vtile_sizes(nvecs: int, vmax: int, granularity: int = 64):
    """src/summit/ldscore/gw_ldscore.py:687-697"""
    width = min(nvecs, vmax)
    if width >= granularity:
        width = max(granularity, (width // granularity) * granularity)
    return [min(width, nvecs - s) for s in range(0, nvecs, width)]


@dataclass
class GwResult:
    ldscore: np.ndarray       # (M, B) float64, null-subtracted
    meansq_accum: np.ndarray  # (M, B) raw accumulators
    inv: np.ndarray           # (M,) residual scales
    mc_per_snp_sumsq: np.ndarray | None = None


def gw_ldscore(bed, n_total: int, m_total: int, annot=None, nvecs: int = 10, step_size: int = 1000,
               seed: int = 1, ddof: int = 1, impute: str = "mean", vtiles=None, rows=None,
               rand_dist: str = "rademacher", eps: float = 1e-10, mc: bool = False,
               impute_seed: int | None = None) -> GwResult:
    if annot is None:
        annot = np.ones((m_total, 1), dtype=np.float64)
    annot = np.asarray(annot, dtype=np.float64)
    B = annot.shape[1]
    if impute_seed is None:
        impute_seed = int((np.uint64(seed) ^ np.uint64(0xA24BAED4963EE407)) & np.uint64(0xFFFFFFFFFFFFFFFF))
    N = n_total if rows is None else len(rows)
    blocks = [(s, min(m_total, s + step_size)) for s in range(0, m_total, step_size)]
    if vtiles is None:
        vtiles = [nvecs]
    assert sum(vtiles) == nvecs
    Xblk = [read_block_standardized(bed, n_total, s, e, rows, ddof, impute, impute_seed) for s, e in blocks]
    inv_all = np.concatenate([inv_resvar(X, eps) for X in Xblk])
    meansq = np.zeros((m_total, B), dtype=np.float64)
    mcps = np.zeros((m_total, B), dtype=np.float64) if mc else None
    N_denom = N - ddof + 1
    v_start = 0
    for Vt in vtiles:
        Xz = np.zeros((N, B * Vt), dtype=np.float64, order="F")
        for (s, e), X in zip(blocks, Xblk):
            Z = probe_block(seed, s, v_start, e - s, Vt, rand_dist)
            Xz += phase1_chunk(X, annot[s:e], inv_all[s:e], Z)
        for (s, e), X in zip(blocks, Xblk):
            phase2_accum(X, inv_all[s:e], Vt, Xz, meansq, N_denom, s, mc_per_snp_sumsq=mcps)
        v_start += Vt
    nsnps_bin = annot.sum(axis=0)
    ld = meansq / float(nvecs) - (nsnps_bin / float(N - ddof))[None, :]
    return GwResult(ld, meansq, inv_all, mcps)


def exact_ld(bed, n_total: int, m_total: int, rows=None, ddof: int = 1, impute: str = "mean",
             impute_seed: int = 0, eps: float = 1e-10, annot=None):
    """Exact LD scores sum_i r_ij^2 a_ik - M_k/(N-1) from the dense correlation
    (tests/test_pgen_gwld.py:732-739)."""
    X = read_block_standardized(bed, n_total, 0, m_total, rows, ddof, impute, impute_seed)
    N = X.shape[0]
    d = float(N - 1)
    x = X * inv_resvar(X, eps)[None, :]
    corr = (x.T @ x) / d
    r2 = corr * corr
    if annot is None:
        annot = np.ones((m_total, 1))
    annot = np.asarray(annot, dtype=np.float64)
    return r2 @ annot - annot.sum(axis=0)[None, :] / d, corr


# ----------------------------------------------------------------------------
# Exact windowed LD: src/native/winldcore.cpp:215-267 (restandardize),
# :341-372 (panel preparation), :409-478 (unbiased r^2, BP mask), :480-594
# (accumulation), :783-931 (driver); no covariates.
# ----------------------------------------------------------------------------
def restandardize_cols(X):
    """Centre over all N rows, scale to population variance 1 (winldcore.cpp:215-267)."""
    X = np.array(X, dtype=np.float64, order="F")
    N = X.shape[0]
    s = X.sum(axis=0)
    ss = (X * X).sum(axis=0)
    mean = s / N
    m2 = ss - s * s / N
    m2 = np.where((m2 < 0) & (m2 > -1e-12), 0.0, m2)
    inv = np.ones_like(mean)
    ok = (m2 > 0) & np.isfinite(m2)
    var = np.where(ok, m2 / N, 0.0)
    ok &= var > 0
    inv[ok] = 1.0 / np.sqrt(var[ok])
    return (X - mean[None, :]) * inv[None, :]


def windowed_ld_chr(bed, n_total: int, chr_start: int, chr_end: int, bp, annot_chr, ld_wind_kb: float = None,
                    rows=None, impute: str = "mean", impute_seed: int = 0, window_snps: int = 0, C=None, R=None):
    """L[j, b] = sum_{i in window(j)} (r_ij^2 - (1 - r_ij^2)/(corr_dim - 1)) a[i, b].

    With covariates C (N, p) and R (p, N) every panel is projected, X <- X - C (R X),
    and restandardized (winldcore.cpp:341-372); corr_dim = N - p, else N - 1
    (winldcore.cpp:848-851)."""
    X = read_block_standardized(bed, n_total, chr_start, chr_end, rows, 0, impute, impute_seed)
    p = 0
    if C is not None and R is not None:
        C = np.asarray(C, dtype=np.float64)
        R = np.asarray(R, dtype=np.float64)
        p = C.shape[1]
        X = restandardize_cols(X - C @ (R @ X))
    elif impute == "mean":
        X = restandardize_cols(X)
    N = X.shape[0]
    corr_dim = N - (p if p > 0 else 1)
    r = (X.T @ X) / N
    r2 = r * r
    r2u = r2 - (1.0 - r2) / (corr_dim - 1)
    m = chr_end - chr_start
    if window_snps > 0:
        idx = np.arange(m)
        mask = np.abs(idx[:, None] - idx[None, :]) <= window_snps
    else:
        bpv = np.asarray(bp, dtype=np.float64)
        mask = ~(np.abs(bpv[:, None] - bpv[None, :]) > ld_wind_kb * 1000.0)
    r2u = np.where(mask, r2u, 0.0)
    return r2u @ np.asarray(annot_chr, dtype=np.float64)


# ----------------------------------------------------------------------------
# SUMMIT-GxE directional scores, Python backend of GenomewideEnvLDScore:
# probes src/summit/ldscore/gwe_ldscore.py:515-529, :2484-2505; features
# :2164-2191, :2323-2375; sources :2561-2579, :3679-3709; targets and score
# accumulation :2581-2586, :3738-3759; final division by nvecs :3826.
# G is the mean-imputed standardized genotype block; C_int an orthonormal,
# centred covariate basis (N x p); env the environment vector.
# ----------------------------------------------------------------------------
def gxe_probe_block(root: int, blk_start: int, probe_ids, L: int, dist: str = "rademacher",
                    fp32: bool = False) -> np.ndarray:
    """(L, V) probes from NumPy Philox(_make_seed(root, blk_start, probe_id))
    (src/summit/ldscore/gwe_ldscore.py:2484-2505): +-1 from integers {0, 1},
    or standard normals (rescaled to norm sqrt(L) for "spherical"); fp32
    rounds to float32 as a float32 run stores them."""
    Z = np.empty((L, len(probe_ids)), dtype=np.float64, order="F")
    for c, pid in enumerate(probe_ids):
        rng = np.random.Generator(np.random.Philox(make_seed(root, blk_start, int(pid))))
        if dist == "rademacher":
            Z[:, c] = 2.0 * rng.integers(0, 2, size=L, dtype=np.int8).astype(np.float64) - 1.0
        else:
            v = rng.standard_normal(size=L)
            if dist == "spherical":
                nrm = float(np.linalg.norm(v))
                v = v * (np.sqrt(float(L)) / nrm) if nrm > 0.0 else np.zeros_like(v)
            Z[:, c] = v
    return Z.astype(np.float32).astype(np.float64) if fp32 else Z


def gxe_project_center(M, C_int):
    M = np.array(M, dtype=np.float64, order="F")
    if C_int is not None and C_int.shape[1] > 0:
        M -= C_int @ (C_int.T @ M)
    M -= M.mean(axis=0, keepdims=True)
    return M


def gxe_scores(bed, n_total: int, m_total: int, env, C_int, annot=None, nvecs: int = 8, seed: int = 1,
               step_size: int = 1000, vtiles=None, ddof: int = 1, impute: str = "mean", rows=None,
               probe_offset: int = 0, eps_var: float = 1e-10, rand_dist: str = "rademacher", probe_fp32: bool = False):
    env = np.asarray(env, dtype=np.float64)
    C_int = np.zeros((env.size, 0)) if C_int is None else np.asarray(C_int, dtype=np.float64)
    if annot is None:
        annot = np.ones((m_total, 1))
    annot = np.asarray(annot, dtype=np.float64)
    K = annot.shape[1]
    G = read_block_standardized(bed, n_total, 0, m_total, rows, ddof, impute, 0)
    N = G.shape[0]
    df = float(N - C_int.shape[1] - 1)
    X = gxe_project_center(G, C_int)
    W = gxe_project_center(env[:, None] * G, C_int)
    inv_x = 1.0 / np.sqrt(np.sum(X * X, axis=0) / df)
    inv_w = 1.0 / np.sqrt(np.sum(W * W, axis=0) / df)
    X *= inv_x[None, :]
    W *= inv_w[None, :]
    blocks = [(s, min(m_total, s + step_size)) for s in range(0, m_total, step_size)]
    vtiles = [nvecs] if vtiles is None else list(vtiles)
    acc = {k: np.zeros((m_total, K)) for k in ("xx", "xw", "wx", "ww")}
    v0 = 0
    for Vt in vtiles:
        kt = K * Vt
        src = np.zeros((N, 2 * kt))
        for s, e in blocks:
            Z = gxe_probe_block(seed, s, [probe_offset + v0 + v for v in range(Vt)], e - s, rand_dist, probe_fp32)
            sq = np.sqrt(np.maximum(annot[s:e], 0.0))
            for k in range(K):
                wz = sq[:, k:k + 1] * Z
                src[:, k * Vt:(k + 1) * Vt] += X[:, s:e] @ wz
                src[:, kt + k * Vt:kt + (k + 1) * Vt] += W[:, s:e] @ wz
        wx_ = X.T @ src
        ww_ = W.T @ src
        for name, work, off in (("xx", wx_, 0), ("xw", wx_, kt), ("wx", ww_, 0), ("ww", ww_, kt)):
            for k in range(K):
                seg = work[:, off + k * Vt:off + (k + 1) * Vt]
                acc[name][:, k] += np.sum(seg * seg, axis=1) / (df * df)
        v0 += Vt
    out = {k: v / float(nvecs) for k, v in acc.items()}
    out["inv_x"] = inv_x
    out["inv_w"] = inv_w
    return out


# ----------------------------------------------------------------------------
# GxE DirectContext, the per-block native GxE backend (src/native/gxeldcore.cpp:
# 7086-8877): decode :8656-8766, features :7349-7416 and :8033-8243, sources
# :7418-7576, projected panels :7578-7680, targets :7682-7826 and :8341-8470,
# phenotype scores :7839-8029.  q is the orthonormal (N, q) design basis spanning the
# intercept and the environment, P = I - q q^T, rank = N - q.
# ----------------------------------------------------------------------------
def direct_decode(bed: np.ndarray, n_total: int, blk_start: int, blk_end: int, rows=None, ddof: int = 1):
    """DirectContext::decode_block_into (gxeldcore.cpp:8674-8766): X (N, L) F-order with
    x = (mean - g) inv_sd at observed calls (g the compact count 00->0, 10->1, 11->2,
    so x is the standardized A1 dosage) and 0 at missing calls; mean over the observed
    calls, inv_sd = sqrt((N - ddof) / m2) with m2 = sum g^2 - (sum g)^2 / nobs over
    observed calls and N the selected sample count (1 when m2 <= 0 or N - ddof <= 0;
    m2 in (-1e-12, 0) reads 0).  Returns (X, missing-call count)."""
    codes, nobs, sm, ss = decode_block(bed, n_total, blk_start, blk_end, rows)
    L, N = codes.shape
    sf, ssf, dn = sm.astype(np.float64), ss.astype(np.float64), nobs.astype(np.float64)
    with np.errstate(divide="ignore", invalid="ignore"):
        mean = np.where(nobs > 0, sf / np.where(nobs > 0, dn, 1.0), 0.0)
        m2 = np.where(nobs > 0, ssf - sf * sf / np.where(nobs > 0, dn, 1.0), 0.0)
    m2 = np.where((m2 < 0.0) & (m2 > -1e-12), 0.0, m2)
    denom = N - ddof
    inv = np.ones(L)
    ok = (denom > 0) & (m2 > 0.0)
    inv[ok] = np.sqrt(float(denom) / m2[ok])
    val = np.array([0.0, 0.0, 1.0, 2.0])[codes]
    X = np.where(codes == CODE_MISSING, 0.0, (mean[:, None] - val) * inv[:, None])
    return np.asfortranarray(X.T), int((N - nobs).sum())


def direct_features(X: np.ndarray, env, q) -> dict:
    """Projected, scaled GxE features of a decoded block, in dense form: x = P X_j,
    w = P (e . X_j), scale = sqrt(rank / ||.||^2), and the feature moments
    norm = ||scaled||^2 / rank, diag_nxe = ||e . scaled||^2 / rank,
    corr_xw = <x, w> / rank (the quantities compute_feature_moments_from_genotype,
    gxeldcore.cpp:8033-8243, forms from Q^T E^p X moments).  The projected
    features x, w are returned scaled."""
    X = np.asarray(X, dtype=np.float64)
    env = np.asarray(env, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    N = X.shape[0]
    rank = N - q.shape[1]
    x = X - q @ (q.T @ X)
    ex = env[:, None] * X
    w = ex - q @ (q.T @ ex)
    ssx = np.sum(x * x, axis=0)
    ssw = np.sum(w * w, axis=0)
    scale_x = np.sqrt(rank / ssx)
    scale_w = np.sqrt(rank / ssw)
    x = x * scale_x
    w = w * scale_w
    e2 = (env * env)[:, None]
    return {"x": x, "w": w, "scale_x": scale_x, "scale_w": scale_w,
            "norm_x": np.sum(x * x, axis=0) / rank, "norm_w": np.sum(w * w, axis=0) / rank,
            "diag_nxe_x": np.sum(e2 * x * x, axis=0) / rank, "diag_nxe_w": np.sum(e2 * w * w, axis=0) / rank,
            "corr_xw": np.sum(x * w, axis=0) / rank, "ssx": ssx, "ssw": ssw, "rank": rank}


def direct_source(X, env, q, scale_x, scale_w, sqrt_annotation, probes):
    """DirectContext::source_block (gxeldcore.cpp:7418-7576): the fused source
    [P X diag(a sx) Z, P (e . X diag(a sw) Z)], (N, 2V)."""
    X = np.asarray(X, dtype=np.float64)
    env = np.asarray(env, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    Z = np.asarray(probes, dtype=np.float64)
    a = np.asarray(sqrt_annotation, dtype=np.float64)[:, None]
    sx = X @ (Z * a * np.asarray(scale_x, dtype=np.float64)[:, None])
    sw = env[:, None] * (X @ (Z * a * np.asarray(scale_w, dtype=np.float64)[:, None]))
    S = np.hstack([sx, sw])
    return S - q @ (q.T @ S)


def direct_leakage(q, S) -> float:
    """||q^T S||_F / ||S||_F (projected_source_leakage_from_coefficients,
    gxeldcore.cpp:8244-8287)."""
    S = np.asarray(S, dtype=np.float64)
    c = np.asarray(q, dtype=np.float64).T @ S
    return float(np.sqrt(np.sum(c * c) / max(np.sum(S * S), np.finfo(np.float64).tiny)))


def direct_target(X, env, q, scale_x, scale_w, S):
    """DirectContext::target_block (gxeldcore.cpp:7682-7826): with R = P S,
    work_x = diag(sx) X^T R and work_w = diag(sw) X^T (e . R); target_projected_block
    and phenotype scores (:8341-8470, :7839-8029) are the same products of an already
    projected panel."""
    X = np.asarray(X, dtype=np.float64)
    env = np.asarray(env, dtype=np.float64)
    q = np.asarray(q, dtype=np.float64)
    S = np.asarray(S, dtype=np.float64)
    R = S - q @ (q.T @ S)
    wx = (X.T @ R) * np.asarray(scale_x, dtype=np.float64)[:, None]
    ww = (X.T @ (env[:, None] * R)) * np.asarray(scale_w, dtype=np.float64)[:, None]
    return wx, ww


# ----------------------------------------------------------------------------
# Exact integer products on masked genotype counts (missing -> 0)
# ----------------------------------------------------------------------------
def masked_geno(codes) -> np.ndarray:
    return np.array([0, 0, 1, 2], dtype=np.int64)[codes]


def mailman_segment_size(n: int) -> int:
    """src/native/common/genotype.cpp:186-200 (no environment override)."""
    return int(_lib().or_mailman_segment_size(int(n)))


def mailman_gz(packed: np.ndarray, n: int, Z, threads: int = 0) -> np.ndarray:
    """U (n, b) int64 = G.Z through the Mailman post-multiply
    (src/native/common/mailman.hpp:213-270) over packed SNP-major BED rows
    (M, stride) with missing calls as 0; Z (M, b) int8."""
    p = np.ascontiguousarray(packed, dtype=np.uint8)
    z = np.ascontiguousarray(Z, dtype=np.int8)
    m, b = z.shape
    out = np.empty((n, b), dtype=np.int64)
    _lib().or_mailman_gz(p.ctypes.data, p.shape[1], n, m, z.ctypes.data, b, out.ctypes.data, int(threads))
    return out


def mailman_gtu(packed: np.ndarray, n: int, U, threads: int = 0) -> np.ndarray:
    """W (M, b) int64 = G^T U through the Mailman pre-multiply
    (src/native/common/mailman.hpp:17-78); U (n, b) int64."""
    p = np.ascontiguousarray(packed, dtype=np.uint8)
    u = np.ascontiguousarray(U, dtype=np.int64)
    b = u.shape[1]
    m = p.shape[0]
    out = np.empty((m, b), dtype=np.int64)
    _lib().or_mailman_gtu(p.ctypes.data, p.shape[1], n, m, u.ctypes.data, b, out.ctypes.data, int(threads))
    return out


def gz_int(g_masked, Z) -> np.ndarray:
    """U[n, b] = sum_i g[i, n] * Z[i, b] (int64)."""
    return np.asarray(g_masked, dtype=np.int64).T @ np.asarray(Z, dtype=np.int64)


def gtu_int(g_masked, U) -> np.ndarray:
    """W[i, b] = sum_n g[i, n] * U[n, b] (int64)."""
    return np.asarray(g_masked, dtype=np.int64) @ np.asarray(U, dtype=np.int64)
