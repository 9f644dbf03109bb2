"""Drop-in mirror of the reference's ``summit.gwldcore`` native module (hot-path exports).

Same function names, keyword arguments, array contract and error texts as the
nanobind module (src/native/gwldcore.cpp:3656-3829); the work runs in
libsummit_gpe on the GPU.  Differences a caller can see:

* The BED file is made resident in HBM on first use (an LRU cache keyed by
  file identity and sample selection replaces the mmap cache of
  src/native/common/genotype.cpp:350-410); ``prefetch_bed_block`` loads it.
* Arithmetic is FP64 whatever the array dtype; float32 arrays receive the
  FP64 results rounded to float32.  Gaussian / spherical probes are drawn in
  the arrays' precision, as the reference's std::normal_distribution<T> does.
* Arrays may be NumPy arrays (host, copied to and from the GPU per call) or
  torch tensors already on the GPU (no copies: ``Xz2d`` as a contiguous (Q, N)
  tensor for the column-major layout, ``meansq_accum`` as (M, B)).
* Covariate projection (``C``/``R``), column sums and ``resvar_gram`` run on
  the GPU in the library's own FP64 kernels (gpe_project_f64, gpe_col_sums_f64,
  the phase-2 contraction); the Cholesky factor of the small p x p
  ``resvar_gram`` is taken on the host in FP64.
* ``precompute_residual_variances_bed(compute_mu22=True)`` forms the per-sample
  sums S of the fourth-moment statistics on the GPU (gpe_sq_rowsums_f64, at
  most 64 covariates) and the B x B ratios on the host in FP64;
  ``compute_block_corrections_binary`` is host FP64.
"""
from __future__ import annotations

import os
import threading
from collections import OrderedDict

import numpy as np
import torch

from . import engine

_LOCK = threading.RLock()
_GENO_CACHE: "OrderedDict[tuple, engine.Genotype]" = OrderedDict()
_GENO_CACHE_CAP = 2
_LINES_CACHE: dict = {}
_STATS = {"hits": 0, "misses": 0, "evictions": 0}
_VERBOSE = False
_THREADS = max(1, len(os.sched_getaffinity(0)))

EXTREME_LD_WARN = float(os.environ.get("SUMMIT_P2_EXTREME_LD_WARN", os.environ.get("SUMMIT_P2_ABS_LD_MAX", "1e10")))
DROP_EXTREME = os.environ.get("SUMMIT_P2_DROP_EXTREME", "1").strip().lower() not in {"0", "false", "no"}
LIMBS = int(os.environ.get("SUMMIT_GPE_LIMBS", "5"))


# --------------------------------------------------------------------------- helpers
def _identity(path: str) -> tuple:
    st = os.stat(path)
    return (os.path.realpath(path), st.st_dev, st.st_ino, st.st_size, st.st_mtime_ns)


def _count_lines(path: str) -> int:
    """count_lines_cached (src/native/common/genotype.cpp:76-97)."""
    key = _identity(path)
    n = _LINES_CACHE.get(key)
    if n is None:
        with open(path, "rb") as f:
            n = 0
            last = b"\n"
            for chunk in iter(lambda: f.read(1 << 20), b""):
                n += chunk.count(b"\n")
                last = chunk[-1:]
            if last != b"\n":
                n += 1
        _LINES_CACHE[key] = n
    return n


def _parse_row_sel(row_sel, n_total: int):
    """parse_row_sel_nb (src/native/common/nb_utils.hpp:94-149)."""
    if row_sel is None:
        return None
    a = np.asarray(row_sel)
    if a.ndim != 1:
        raise RuntimeError("row_sel must be a 1D CPU array of int32 or int64 indices")
    if a.dtype not in (np.int32, np.int64):
        raise RuntimeError("row_sel must have dtype int32 or int64")
    a = a.astype(np.int64)
    if a.size and (a.min() < 0 or a.max() >= n_total or a.max() > np.iinfo(np.int32).max):
        raise RuntimeError("row_sel contains an out-of-range sample index")
    return np.ascontiguousarray(a)


def _parse_impute(mode: str) -> str:
    if mode not in ("mean", "hwe"):
        raise RuntimeError("impute_mode must be 'mean' or 'hwe'")
    return mode


def _geno(bed_prefix: str, fam_path: str, row_sel, ddof: int, impute_mode: str, impute_seed) -> tuple:
    bed_path = bed_prefix + ".bed"
    bim_path = bed_prefix + ".bim"
    n_total = _count_lines(fam_path)
    if n_total <= 0:
        raise RuntimeError("FAM has zero rows: " + fam_path)
    m_total = _count_lines(bim_path)
    rows = _parse_row_sel(row_sel, n_total)
    key = (_identity(bed_path), n_total, m_total, None if rows is None else rows.tobytes())
    seed = 0 if impute_seed is None else int(impute_seed)
    with _LOCK:
        g = _GENO_CACHE.get(key)
        if g is None:
            _STATS["misses"] += 1
            if not os.path.exists(bed_path):
                raise RuntimeError("Failed to open bed: " + bed_path)
            raw = np.memmap(bed_path, dtype=np.uint8, mode="r")
            if raw.size < 3:
                raise RuntimeError("stat failed or BED too small: " + bed_path)
            if not (raw[0] == 0x6C and raw[1] == 0x1B and raw[2] == 0x01):
                raise RuntimeError("Invalid BED magic/mode (expected 6C 1B 01) for: " + bed_path)
            stride = (n_total + 3) // 4
            if raw.size < 3 + stride * m_total:
                raise RuntimeError("BED file too small for requested block: " + bed_path)
            body = raw[3:3 + stride * m_total].reshape(m_total, stride)
            g = engine.Genotype.from_bed_bytes(body, n_total, m_total, rows=rows)
            del raw, body
            _GENO_CACHE[key] = g
            while len(_GENO_CACHE) > _GENO_CACHE_CAP:
                _, old = _GENO_CACHE.popitem(last=False)
                old.close()
                _STATS["evictions"] += 1
        else:
            _STATS["hits"] += 1
            _GENO_CACHE.move_to_end(key)
        want = (int(ddof), _parse_impute(impute_mode), seed if impute_mode == "hwe" else 0)
        if g.prepared_with != want:
            g.prepare(ddof=int(ddof), impute=impute_mode, impute_seed=seed, snp_offset=0, global_rows=rows)
            g.prepared_with = want
    return g, n_total, m_total


def _dev():
    return engine.default_device()


def _to_dev(a, dtype=torch.float64):
    if a is None:
        return None
    if isinstance(a, torch.Tensor):
        return a.to(device=_dev(), dtype=dtype)
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a), dtype=np.float64 if dtype == torch.float64 else None)).to(_dev(), dtype)


def _xz_colmajor_dev(x):
    """Fortran (N, Q) array or (Q, N) device tensor -> contiguous (Q, N) FP64 device tensor."""
    if isinstance(x, torch.Tensor):
        if x.is_cuda and x.dtype == torch.float64 and x.is_contiguous():
            return x, False
        return x.to(device=_dev(), dtype=torch.float64).contiguous(), True
    a = np.asarray(x)
    # an Xz2d the caller allocated with np.empty may hold signalling-NaN bit
    # patterns where phase 1 is about to overwrite (bin_init_mask); converting
    # them is harmless and not worth a warning
    with np.errstate(invalid="ignore"):
        return torch.from_numpy(np.ascontiguousarray(a.T, dtype=np.float64)).to(_dev()), True


def _write_back_colmajor(dst, dev):
    if isinstance(dst, torch.Tensor):
        if dst.data_ptr() != dev.data_ptr():
            dst.copy_(dev.to(dst.dtype))
        return
    dst[...] = dev.cpu().numpy().T.astype(dst.dtype, copy=False)


def _fresh(x: torch.Tensor) -> torch.Tensor:
    """A new contiguous FP64 copy of a device matrix."""
    out = torch.empty(x.shape, dtype=torch.float64, device=x.device)
    out.copy_(x)
    return out


def _project_dev(x_qn: torch.Tensor, C, R) -> torch.Tensor:
    """(I - C R) x for x column-major (N, Q) held as (Q, N), into a new tensor
    (gpe_project_f64: FP64 coefficients C^T-side reduction, then the update)."""
    return engine.project(_fresh(x_qn), _to_dev(C), _to_dev(R))


def _project_t_dev(x_qn: torch.Tensor, C, R) -> torch.Tensor:
    """x - R^T (C^T x): the transpose projection that carries phase 2's
    projected genotypes onto X.Z (G_p^T x = G^T (I - CR)^T x), into a new tensor."""
    Cd = _to_dev(C)
    Rd = _to_dev(R)
    return engine.project(_fresh(x_qn), Rd.T, Cd.T)


# --------------------------------------------------------------------------- plumbing exports
def set_verbose(enabled: bool) -> None:
    global _VERBOSE
    _VERBOSE = bool(enabled)


def set_num_threads(n: int) -> None:
    global _THREADS
    _THREADS = max(1, int(n))


def get_max_threads() -> int:
    return _THREADS


def prefetch_bed_block(bed_prefix, fam_path, blk_start, blk_end, ahead_blocks=1, impute_mode="mean", ddof=1,
                       impute_seed=None) -> None:
    """Makes the BED resident in HBM (the whole file on first use)."""
    _geno(bed_prefix, fam_path, None, ddof, impute_mode, impute_seed)


def _bed_mapping_cache_info() -> dict:
    return {"entries": len(_GENO_CACHE), "capacity": _GENO_CACHE_CAP, **_STATS}


def _clear_bed_mapping_cache() -> None:
    with _LOCK:
        for g in _GENO_CACHE.values():
            g.close()
        _GENO_CACHE.clear()
        for k in _STATS:
            _STATS[k] = 0


def clear_phase1_native_scratch() -> None:
    pass


def clear_phase2_native_scratch() -> None:
    pass


def clear_gwld_genotype_scratch() -> None:
    pass


def clear_gwld_native_scratch() -> None:
    pass


def clear_phase1_csr_cache() -> None:
    pass


# --------------------------------------------------------------------------- phase 1
def _bin_nnz(annot: np.ndarray, inv: np.ndarray) -> np.ndarray:
    """Entries the phase-1 CSR keeps per bin (src/native/gwldcore.cpp:1002-1063)."""
    a = np.asarray(annot, dtype=np.float64)
    iv = np.asarray(inv, dtype=np.float64)[:, None]
    with np.errstate(invalid="ignore", over="ignore"):
        sc = iv * np.sqrt(np.where(a > 0, a, 0.0))
    ok = (a != 0) & (a > 0) & np.isfinite(a) & np.isfinite(iv) & (iv > 0) & np.isfinite(sc)
    return ok.sum(axis=0)


def _phase1(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, annot_blk, inv_right, v_start, v_count,
            rand_dist, seed, Xz2d_chunk, bin_init_mask, project_right, C, R, impute_mode, impute_seed, resvar_gram,
            inv_out, resvar_eps, rowmajor):
    g, n_total, m_total = _geno(bed_prefix, fam_path, row_sel, ddof, impute_mode, impute_seed)
    blk_start, blk_end = int(blk_start), int(blk_end)
    if blk_end > m_total:
        raise RuntimeError("blk_end exceeds #SNPs in BIM")
    L = max(0, blk_end - blk_start)
    if L == 0:
        return
    if project_right and C is not None and R is not None and inv_out is not None:
        raise RuntimeError("Fused residual variance is incompatible with project_right=True")
    annot = np.asarray(annot_blk.cpu().numpy() if isinstance(annot_blk, torch.Tensor) else annot_blk,
                       dtype=np.float64)
    if annot.ndim != 2 or annot.shape[0] != L:
        raise RuntimeError("L mismatch")
    B = annot.shape[1]
    N = g.n
    if inv_out is not None:
        if len(inv_out) != L:
            raise RuntimeError("inv_out shape mismatch in phase1 fused residual variance")
        inv_d = residual_inv_block(g, blk_start, blk_end, float(resvar_eps), C, R, resvar_gram)
        if isinstance(inv_out, torch.Tensor):
            inv_out.copy_(inv_d.to(inv_out.dtype))
        else:
            inv_out[...] = inv_d.cpu().numpy().astype(inv_out.dtype, copy=False)
        inv_np = inv_d.cpu().numpy()
    else:
        if inv_right is None or len(inv_right) != L:
            raise RuntimeError("inv_right shape mismatch in phase1")
        inv_np = np.asarray(inv_right.cpu().numpy() if isinstance(inv_right, torch.Tensor) else inv_right,
                            dtype=np.float64)
        inv_d = _to_dev(inv_np)
    shape_q = B * int(v_count)
    if rowmajor:
        if tuple(Xz2d_chunk.shape) != (N, shape_q):
            raise RuntimeError("Xz2d_chunk shape must be (N, B*v_count)")
    else:
        ok = (tuple(Xz2d_chunk.shape) == (N, shape_q)) if not isinstance(Xz2d_chunk, torch.Tensor) else \
            (tuple(Xz2d_chunk.shape) in ((shape_q, N),))
        if not ok:
            raise RuntimeError("Xz2d_chunk shape must be (N, B*v_count)")
    if seed is None:
        seed = int.from_bytes(os.urandom(8), "little")
    nnz = _bin_nnz(annot, inv_np)
    modes = np.full(B, 2, dtype=np.uint8)
    if bin_init_mask is not None:
        if len(bin_init_mask) != B:
            raise RuntimeError("bin_init_mask shape mismatch in phase1")
        for k in range(B):
            if nnz[k] > 0:
                modes[k] = 2 if bin_init_mask[k] else 1
                bin_init_mask[k] = 1
    projecting = bool(project_right) and C is not None and R is not None
    annot_d = _to_dev(annot)
    # the reference draws its probes in the precision of the caller's arrays
    # (std::normal_distribution<T>, src/native/gwldcore.cpp:1694)
    fp32 = str(getattr(Xz2d_chunk, "dtype", "")).endswith("float32")
    if rowmajor or projecting:
        contrib = torch.zeros((shape_q, N), dtype=torch.float64, device=_dev())
        engine.phase1_xz(g, blk_start, blk_end, L, annot_d, inv_d, int(v_start), int(v_count), int(seed), contrib,
                         rand_dist=rand_dist, limbs=LIMBS, probe_fp32=fp32)
        if projecting:
            contrib = _project_dev(contrib, C, R)
        if rowmajor:
            if isinstance(Xz2d_chunk, torch.Tensor):
                Xz2d_chunk += contrib.T.to(Xz2d_chunk.dtype)
            else:
                Xz2d_chunk += contrib.cpu().numpy().T.astype(Xz2d_chunk.dtype, copy=False)
            return
        xz, _ = _xz_colmajor_dev(Xz2d_chunk)
        for k in range(B):
            sl = slice(k * v_count, (k + 1) * v_count)
            if modes[k] == 1:
                xz[sl] = contrib[sl]
            elif nnz[k] > 0:
                xz[sl] += contrib[sl]
        _write_back_colmajor(Xz2d_chunk, xz)
        return
    xz, _ = _xz_colmajor_dev(Xz2d_chunk)
    mode_d = torch.from_numpy(modes).to(_dev())
    engine.phase1_xz(g, blk_start, blk_end, L, annot_d, inv_d, int(v_start), int(v_count), int(seed), xz,
                     rand_dist=rand_dist, bin_mode=mode_d, limbs=LIMBS, probe_fp32=fp32)
    _write_back_colmajor(Xz2d_chunk, xz)


def phase1_compute_Xz_bed_chunk(bed_prefix, fam_path, blk_start, blk_end, row_sel=None, ddof=1, *, annot_blk,
                                inv_right, v_start, v_count, kmax_hint, rand_dist="rademacher", seed=None,
                                Xz2d_chunk, bin_init_mask=None, project_right=False, C=None, R=None,
                                impute_mode="hwe", impute_seed=None, resvar_gram=None, inv_out=None,
                                resvar_eps=1e-10):
    """Phase 1, Xz (+)= X_blk diag(inv*sqrt(a)) Z for one SNP block
    (src/native/gwldcore.cpp:1582-1923; Xz2d_chunk Fortran (N, B*v_count))."""
    _phase1(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, annot_blk, inv_right, v_start, v_count,
            rand_dist, seed, Xz2d_chunk, bin_init_mask, project_right, C, R, impute_mode, impute_seed, resvar_gram,
            inv_out, resvar_eps, rowmajor=False)


def phase1_compute_Xz_bed_chunk_rowmajor(bed_prefix, fam_path, blk_start, blk_end, row_sel=None, ddof=1, *,
                                         annot_blk, inv_right, v_start, v_count, kmax_hint,
                                         rand_dist="rademacher", seed=None, Xz2d_chunk, project_right=False, C=None,
                                         R=None, impute_mode="hwe", impute_seed=None, resvar_gram=None,
                                         inv_out=None, resvar_eps=1e-10):
    """Row-major variant (src/native/gwldcore.cpp:1925-2130): accumulates into C-contiguous (N, B*v_count)."""
    _phase1(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, annot_blk, inv_right, v_start, v_count,
            rand_dist, seed, Xz2d_chunk, None, project_right, C, R, impute_mode, impute_seed, resvar_gram,
            inv_out, resvar_eps, rowmajor=True)


# --------------------------------------------------------------------------- residual variances
def _residual_sums(g: engine.Genotype, s: int, e: int, C=None, R=None, gram=None):
    """(ss, ss_proj, df) for SNPs [s, e): ||x_j||^2 of the standardized working
    genotypes, ||P x_j||^2 with P x = x - C (R x) (= ss without covariates) and
    the residual degrees of freedom N - p - 1 (N - 1 without covariates).

    ||P x||^2 = ||x||^2 + q - 2h with h = ||R x||^2 and q = h or (R x)^T G (R x);
    both are sums of squares of phase-2 products against transformed covariate
    rows, so they run on the packed-contraction kernel (G = L L^T, q = ||L^T R x||^2)."""
    if (C is None) != (R is None):
        raise RuntimeError("C and R must be provided together for fused residual variances")
    N = g.n
    # plain ||x_j||^2 from the integer moments: inv0 = 1/sqrt(ss/(N-1)) -> ss
    inv0 = engine.resvar_inv(g, s, e, 0.0).cpu().numpy()
    with np.errstate(divide="ignore", invalid="ignore"):
        ss = float(N - 1) / (inv0 * inv0)
    ss = np.where(np.isfinite(ss), ss, 0.0)
    if C is None:
        return ss, ss, float(N - 1)
    Rd = _to_dev(R)
    p = Rd.shape[0]
    if tuple(_to_dev(C).shape) != (N, p) or Rd.shape[1] != N:
        raise RuntimeError("C/R shape mismatch in fused residual variances")
    if N - p <= 1:
        raise RuntimeError("N_eff must be > 1 in fused residual variances")
    ones = torch.ones(e - s, dtype=torch.float64, device=_dev())

    def sumsq_rows(Mrows: torch.Tensor) -> np.ndarray:
        acc = torch.zeros((g.m, 1), dtype=torch.float64, device=_dev())
        engine.phase2_accum(g, s, e, ones, Mrows.shape[0], Mrows.contiguous(), acc, n_denom=2,
                            extreme_warn=0.0, limbs=max(LIMBS, 6))
        return acc[s:e, 0].cpu().numpy()

    h = sumsq_rows(Rd)
    if gram is not None:
        # q = (R x)^T G (R x) = ||L^T R x||^2 with G = L L^T (p x p, host FP64)
        Gh = np.asarray(gram.cpu().numpy() if isinstance(gram, torch.Tensor) else gram, dtype=np.float64)
        Rh = np.asarray(R.cpu().numpy() if isinstance(R, torch.Tensor) else R, dtype=np.float64)
        Lc = np.linalg.cholesky(Gh)
        q = sumsq_rows(_to_dev(Lc.T @ Rh))
    else:
        q = h
    return ss, ss + q - 2.0 * h, float(N - p - 1)


def residual_inv_block(g: engine.Genotype, s: int, e: int, eps: float, C=None, R=None, gram=None) -> torch.Tensor:
    """inv_j = 1/sqrt(max(||P x_j||^2 / df, eps)) (src/native/gwldcore.cpp:3220-3333)."""
    if (C is None) != (R is None):
        raise RuntimeError("C and R must be provided together for fused residual variances")
    if C is None:
        return engine.resvar_inv(g, s, e, eps)
    _, ssp, df = _residual_sums(g, s, e, C, R, gram)
    var = ssp / df
    with np.errstate(invalid="ignore"):
        safe = np.where(np.isnan(var), var, np.maximum(var, eps))
    return _to_dev(1.0 / np.sqrt(safe))


def _mu22(g: engine.Genotype, nsnps: int, step_size: int, annot: np.ndarray, inv: np.ndarray, C, R) -> np.ndarray:
    """mu22[a, b] = (S^T S)[a, b] / (N Msum_a Msum_b) with S[i, k] = sum_j xt_ij^2 a_jk inv_j^2
    (src/native/gwldcore.cpp:3496-3544): S on the GPU (gpe_sq_rowsums_f64, projected
    genotypes when C/R are given), the B x B product and ratios on the host in FP64."""
    N = g.n
    B = annot.shape[1]
    dev = _dev()
    S = torch.zeros((B, N), dtype=torch.float64, device=dev)
    Cd = rx = None
    if C is not None:
        Cd = _to_dev(C)
        if Cd.shape[1] > 64:
            raise RuntimeError("compute_mu22 supports at most 64 covariates")
        rx = engine.xtv(g, _to_dev(R), 0, nsnps, limbs=6)
    inv2 = inv * inv
    # weights a_jk inv_j^2 uploaded in SNP slices of at most 2^22 values
    step = max(1, min(nsnps, (1 << 22) // max(1, B)))
    for s0 in range(0, nsnps, step):
        e0 = min(nsnps, s0 + step)
        w = torch.from_numpy(np.ascontiguousarray((annot[s0:e0] * inv2[s0:e0, None]).T)).to(dev)
        engine.sq_rowsums(g, w, s0, e0, C=Cd, rx=None if rx is None else rx[s0:e0], out=S)
    Sh = S.cpu().numpy()
    msum = annot.sum(axis=0)
    denom = float(N) * msum[:, None] * msum[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        mu22 = np.where(denom != 0.0, (Sh @ Sh.T) / denom, 0.0)
    mu22[~np.isfinite(mu22)] = 0.0
    return np.ascontiguousarray(mu22)


def precompute_residual_variances_bed(bed_prefix, fam_path, nsnps, step_size, row_sel=None, ddof=1, eps=1e-10,
                                      compute_mu22=False, annot_all=None, C=None, R=None, impute_mode="hwe",
                                      impute_seed=None, snp_var_out=None):
    """(inv, mu22 or None) for the first nsnps SNPs (src/native/gwldcore.cpp:3335-3566);
    snp_var_out (nsnps,) receives ||P x_j||^2 / ||x_j||^2 = 1 - R2_j, 0 for a
    monomorphic SNP (:3364-3373, :3491-3493)."""
    g, n_total, m_total = _geno(bed_prefix, fam_path, row_sel, ddof, impute_mode, impute_seed)
    nsnps = int(nsnps)
    if nsnps < 0 or nsnps > m_total:
        raise RuntimeError("nsnps is out of range in precompute_residual_variances_bed")
    if snp_var_out is not None and np.shape(snp_var_out) != (nsnps,):
        raise RuntimeError("snp_var_out shape mismatch in precompute_residual_variances_bed")
    if int(step_size) <= 0:
        raise RuntimeError("step_size must be > 0 in precompute_residual_variances_bed")
    if g.n <= 1:
        raise RuntimeError("Too few rows in precompute_residual_variances_bed")
    have_proj = C is not None and R is not None
    if have_proj:
        p = np.shape(C)[1]
        if np.shape(C)[0] != g.n or np.shape(R) != (p, g.n):
            raise RuntimeError("C/R shape mismatch in precompute_residual_variances_bed")
        if g.n - p <= 1:
            raise RuntimeError("N_eff must be > 1 in precompute_residual_variances_bed")
    else:
        C = R = None
    annot = None
    if compute_mu22:
        if annot_all is None:
            raise RuntimeError("annot_all must be provided when compute_mu22=True")
        annot = np.asarray(annot_all.cpu().numpy() if isinstance(annot_all, torch.Tensor) else annot_all,
                           dtype=np.float64)
        if annot.ndim != 2 or annot.shape[0] != nsnps:
            raise RuntimeError("annot_all shape mismatch in precompute_residual_variances_bed")
        if annot.shape[1] <= 0:
            raise RuntimeError("annot_all must have at least one column")
    if nsnps == 0:
        inv = np.empty(0)
        ss = ssp = np.empty(0)
    elif C is None:
        inv = engine.resvar_inv(g, 0, nsnps, float(eps)).cpu().numpy()
        if snp_var_out is None and annot is None:
            return inv, None
        ss, ssp, _ = _residual_sums(g, 0, nsnps)
    else:
        ss, ssp, df = _residual_sums(g, 0, nsnps, C, R)
        var = ssp / df
        with np.errstate(invalid="ignore", divide="ignore"):
            inv = 1.0 / np.sqrt(np.where(np.isnan(var), var, np.maximum(var, float(eps))))
    if snp_var_out is not None:
        with np.errstate(divide="ignore", invalid="ignore"):
            v = np.where((ss > 0.0) & np.isfinite(ssp), np.minimum(1.0, ssp / np.where(ss > 0.0, ss, 1.0)), 0.0)
        snp_var_out[...] = torch.from_numpy(v) if isinstance(snp_var_out, torch.Tensor) else v
    mu22 = None
    if annot is not None:
        mu22 = _mu22(g, nsnps, int(step_size), annot, inv, C, R) if nsnps > 0 else np.zeros((annot.shape[1],) * 2)
    return inv, mu22


def compute_block_corrections_binary(meansq_raw, mu22, annot_all, N, d):
    """Block-level fourth-moment diagnostics (R2, rho2, bias, delta), each B x B
    (src/native/gwldcore.cpp:3568-3633): over the SNPs i with a_ia != 0,
    S_ab = sum meansq_raw[i, b] and M_a their count; for M_a, M_b > 0 and
    S_ab != 0: R2 = S / (M_a M_b), rho2 = (d^2 R2 - N mu22) / (N (N - 1)),
    bias = S - M_a M_b rho2, delta = mu22 - (1 + 2 rho2); 0 elsewhere."""
    m = np.asarray(meansq_raw, dtype=np.float64)
    if m.ndim != 2:
        raise RuntimeError("meansq_raw must be 2-D")
    M, B = m.shape
    u = np.asarray(mu22, dtype=np.float64)
    if u.shape != (B, B):
        raise RuntimeError("mu22 shape mismatch")
    A = np.asarray(annot_all)
    if A.shape != (M, B):
        raise RuntimeError("annot_all shape mismatch")
    if int(N) <= 1:
        raise RuntimeError("N must be > 1 in compute_block_corrections_binary")
    ind = (A != 0).astype(np.float64)
    Ma = ind.sum(axis=0)
    S = ind.T @ m
    Nf = float(N)
    ok = (Ma[:, None] > 0.0) & (Ma[None, :] > 0.0) & (S != 0.0)
    MM = Ma[:, None] * Ma[None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        R2 = np.where(ok, S / np.where(ok, MM, 1.0), 0.0)
        rho2 = np.where(ok, (float(d) * float(d) * R2 - Nf * u) / (Nf * (Nf - 1.0)), 0.0)
    bias = np.where(ok, S - MM * rho2, 0.0)
    delta = np.where(ok, u - (1.0 + 2.0 * rho2), 0.0)
    return tuple(np.ascontiguousarray(x) for x in (R2, rho2, bias, delta))


# --------------------------------------------------------------------------- phase 2
def _phase2(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, inv_left, tile_nvecs, Xz2d, meansq_accum, C, R,
            N_denom, impute_mode, impute_seed, mc_aggregate_sumsq, mc_per_snp_sumsq, rowmajor, sum_Xz=None):
    g, n_total, m_total = _geno(bed_prefix, fam_path, row_sel, ddof, impute_mode, impute_seed)
    blk_start, blk_end = int(blk_start), int(blk_end)
    if blk_end > m_total:
        raise RuntimeError("blk_end exceeds #SNPs in BIM")
    L = max(0, blk_end - blk_start)
    if L == 0:
        return
    N = g.n
    if inv_left is None or len(inv_left) != L:
        raise RuntimeError("inv_left shape mismatch in phase2_accum_XtXz_bed")
    if rowmajor:
        if isinstance(Xz2d, torch.Tensor):
            xz = Xz2d.to(device=_dev(), dtype=torch.float64).T.contiguous()
        else:
            xz = torch.from_numpy(np.ascontiguousarray(np.asarray(Xz2d, dtype=np.float64).T)).to(_dev())
    else:
        xz, _ = _xz_colmajor_dev(Xz2d)
    if xz.shape[1] != N:
        raise RuntimeError("Xz2d shape mismatch in phase2_accum_XtXz_bed")
    Q = xz.shape[0]
    if tile_nvecs <= 0 or Q % tile_nvecs:
        raise RuntimeError("Xz2d col count must be a multiple of tile_nvecs")
    B = Q // tile_nvecs
    if meansq_accum.shape[1] != B:
        raise RuntimeError("meansq_accum shape mismatch in phase2_accum_XtXz_bed")
    if blk_end > meansq_accum.shape[0]:
        raise RuntimeError("meansq_accum rows smaller than SNP count")
    if mc_aggregate_sumsq is not None and len(mc_aggregate_sumsq) != B:
        raise RuntimeError("mc_aggregate_sumsq length must equal the annotation count")
    if mc_per_snp_sumsq is not None and tuple(mc_per_snp_sumsq.shape) != tuple(meansq_accum.shape):
        raise RuntimeError("mc_per_snp_sumsq shape must match meansq_accum")
    sum_d = None
    if sum_Xz is not None:
        sum_np = sum_Xz.cpu().numpy() if isinstance(sum_Xz, torch.Tensor) else np.asarray(sum_Xz)
        if sum_np.ndim != 1 or sum_np.shape[0] != Q:
            raise RuntimeError("sum_Xz shape mismatch in phase2_accum_XtXz_bed_mailman_rm")
        sum_d = _to_dev(sum_np)
    if C is not None and R is not None:
        xz = _project_t_dev(xz, C, R)
    inv_d = _to_dev(inv_left)
    dev_out = isinstance(meansq_accum, torch.Tensor) and meansq_accum.is_cuda and \
        meansq_accum.dtype == torch.float64 and meansq_accum.is_contiguous()
    if dev_out:
        ms = meansq_accum
        row0 = 0
    else:
        ms = torch.zeros((L, B), dtype=torch.float64, device=_dev())
        row0 = blk_start
    mca = torch.zeros(B, dtype=torch.float64, device=_dev()) if mc_aggregate_sumsq is not None else None
    mcp = torch.zeros((L, B), dtype=torch.float64, device=_dev()) if mc_per_snp_sumsq is not None else None
    if dev_out and mcp is not None and isinstance(mc_per_snp_sumsq, torch.Tensor):
        mcp = mc_per_snp_sumsq
    nonfin, extreme = engine.phase2_accum(
        g, blk_start, blk_end, inv_d, int(tile_nvecs), xz, ms, int(N_denom), mc_aggregate_sumsq=mca,
        mc_per_snp_sumsq=mcp, extreme_warn=EXTREME_LD_WARN, drop_extreme=DROP_EXTREME, limbs=LIMBS,
        meansq_row0=row0, want_guards=True, sum_xz=sum_d)
    if not dev_out:
        upd = ms.cpu().numpy()
        meansq_accum[blk_start:blk_end] += upd.astype(meansq_accum.dtype, copy=False)
        if mcp is not None:
            mc_per_snp_sumsq[blk_start:blk_end] += mcp.cpu().numpy()
    if mca is not None:
        mc_aggregate_sumsq += mca.cpu().numpy() if not isinstance(mc_aggregate_sumsq, torch.Tensor) else mca
    if nonfin or extreme:
        import sys
        print(f"[phase2-watch] accum block [{blk_start}:{blk_end}) nonfinite_dropped={nonfin} extreme={extreme}",
              file=sys.stderr)


def phase2_accum_XtXz_bed(bed_prefix, fam_path, blk_start, blk_end, row_sel=None, ddof=1, *, inv_left, tile_nvecs,
                          Xz2d, meansq_accum, C=None, R=None, N_denom=0, impute_mode="hwe", impute_seed=None,
                          mc_aggregate_sumsq=None, mc_per_snp_sumsq=None):
    """Phase 2, meansq_accum[j, k] += sum_v (x_j . Xz[:, k*V+v])^2 inv_j^2/(N_denom-1)^2
    (src/native/gwldcore.cpp:2329-2649)."""
    _phase2(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, inv_left, tile_nvecs, Xz2d, meansq_accum, C, R,
            N_denom, impute_mode, impute_seed, mc_aggregate_sumsq, mc_per_snp_sumsq, rowmajor=False)


def phase2_accum_XtXz_bed_mailman(bed_prefix, fam_path, blk_start, blk_end, row_sel=None, ddof=1, *, inv_left,
                                  tile_nvecs, Xz2d, meansq_accum, C=None, R=None, N_denom=0, impute_seed=None,
                                  mc_aggregate_sumsq=None, mc_per_snp_sumsq=None):
    """HWE-completed phase 2 (the reference's Mailman path, src/native/gwldcore.cpp:2651-2917):
    the same quantity, computed exactly on int8 MFMA instead of a 3^s lookup table."""
    _phase2(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, inv_left, tile_nvecs, Xz2d, meansq_accum, C, R,
            N_denom, "hwe", impute_seed, mc_aggregate_sumsq, mc_per_snp_sumsq, rowmajor=False)


def phase2_accum_XtXz_bed_mailman_rowmajor(bed_prefix, fam_path, blk_start, blk_end, row_sel=None, ddof=1, *,
                                           inv_left, tile_nvecs, Xz2d, meansq_accum, sum_Xz, N_denom=0,
                                           impute_seed=None, mc_aggregate_sumsq=None, mc_per_snp_sumsq=None):
    """Row-major HWE phase 2 (src/native/gwldcore.cpp:2919-3119): each column c is centred
    with the caller's sum_Xz[c] (the column sums of the projected X.Z the reference's
    driver passes, src/summit/ldscore/gw_ldscore.py:2010-2016)."""
    if sum_Xz is None:
        raise RuntimeError("sum_Xz shape mismatch in phase2_accum_XtXz_bed_mailman_rm")
    _phase2(bed_prefix, fam_path, blk_start, blk_end, row_sel, ddof, inv_left, tile_nvecs, Xz2d, meansq_accum,
            None, None, N_denom, "hwe", impute_seed, mc_aggregate_sumsq, mc_per_snp_sumsq, rowmajor=True,
            sum_Xz=sum_Xz)


# --------------------------------------------------------------------------- column helpers
def _rowmajor_qn_dev(Xz2d):
    """C-contiguous (N, Q) array or tensor -> contiguous (Q, N) FP64 device tensor."""
    if isinstance(Xz2d, torch.Tensor):
        if Xz2d.dim() != 2:
            raise RuntimeError("Xz2d must be a 2D array")
        return Xz2d.to(device=_dev(), dtype=torch.float64).T.contiguous()
    a = np.asarray(Xz2d)
    if a.ndim != 2:
        raise RuntimeError("Xz2d must be a 2D array")
    return torch.from_numpy(np.ascontiguousarray(a.T, dtype=np.float64)).to(_dev())


def compute_col_sums_rowmajor(Xz2d):
    """Column sums of a C-contiguous (N, Q) array (src/native/gwldcore.cpp:3121-3132)."""
    return engine.col_sums(_rowmajor_qn_dev(Xz2d)).cpu().numpy()


def project_rowmajor_inplace_and_col_sums(Xz2d, C=None, R=None):
    """Xz <- (I - C R) Xz for C-contiguous (N, Q), then column sums (src/native/gwldcore.cpp:3134-3171)."""
    x = _rowmajor_qn_dev(Xz2d)
    if C is not None and R is not None:
        x = _project_dev(x, C, R)
        if isinstance(Xz2d, torch.Tensor):
            Xz2d.copy_(x.T.to(Xz2d.dtype))
        else:
            Xz2d[...] = x.cpu().numpy().T.astype(Xz2d.dtype, copy=False)
    return engine.col_sums(x).cpu().numpy()


def project_colmajor_inplace(Xz2d, C=None, R=None) -> None:
    """Xz <- (I - C R) Xz for Fortran (N, Q) (src/native/gwldcore.cpp:3173-3207)."""
    if C is None or R is None:
        return
    xz, _ = _xz_colmajor_dev(Xz2d)
    _write_back_colmajor(Xz2d, _project_dev(xz, C, R))


# --------------------------------------------------------------------------- GRM panel apply
def _panel_shape(panel) -> tuple:
    """(rows N, columns Q) of a Fortran (N, Q) array or a (Q, N) device tensor."""
    if isinstance(panel, torch.Tensor):
        return int(panel.shape[1]), int(panel.shape[0])
    return int(np.shape(panel)[0]), int(np.shape(panel)[1])


def grm_apply_dev(g: engine.Genotype, vin: torch.Tensor, inv_all, nsnps: int, step_size: int, C=None,
                  R=None) -> torch.Tensor:
    """P (1/M) X diag(inv^2) X^T P v for a (Q, N) device panel v (P = I - C R when
    C/R are given), into a new (Q, N) tensor: the device core of apply_grm_bed_panel."""
    projecting = C is not None and R is not None
    if projecting:
        vin = _project_dev(vin, C, R)
    inv = inv_all[:nsnps] if isinstance(inv_all, torch.Tensor) else np.asarray(inv_all)[:nsnps]
    out = engine.apply_grm(g, vin, _to_dev(inv), nsnps, int(step_size), limbs=max(LIMBS, 6))
    if projecting:
        out = _project_dev(out, C, R)
    return out


def _apply_grm(name, bed_prefix, fam_path, nsnps, step_size, row_sel, ddof, inv_all, panel_in, panel_out, C, R,
               impute_mode, impute_seed):
    g, n_total, m_total = _geno(bed_prefix, fam_path, row_sel, ddof, impute_mode, impute_seed)
    nsnps = int(nsnps)
    if nsnps < 0 or nsnps > m_total:
        raise RuntimeError(f"nsnps is out of range in {name}")
    if int(step_size) <= 0:
        raise RuntimeError(f"step_size must be > 0 in {name}")
    if g.n <= 0:
        raise RuntimeError(f"No rows selected in {name}")
    if len(inv_all) < nsnps:
        raise RuntimeError(f"inv_all shape mismatch in {name}")
    (n_in, q_in), (n_out, q_out) = _panel_shape(panel_in), _panel_shape(panel_out)
    if n_in != g.n or n_out != g.n:
        raise RuntimeError(f"panel row mismatch in {name}")
    if q_in != q_out:
        raise RuntimeError(f"panel col mismatch in {name}")
    if C is not None and R is not None:
        p = np.shape(C)[1]
        if np.shape(C)[0] != g.n or tuple(np.shape(R)) != (p, g.n):
            raise RuntimeError(f"C/R shape mismatch in {name}")
    else:
        C = R = None
    vin, _ = _xz_colmajor_dev(panel_in)
    out = grm_apply_dev(g, vin, inv_all, nsnps, int(step_size), C, R)
    _write_back_colmajor(panel_out, out)


def apply_grm_bed_panel(bed_prefix, fam_path, nsnps, step_size, row_sel=None, ddof=1, *, inv_all, panel_in,
                        panel_out, C=None, R=None, impute_mode="hwe", impute_seed=None):
    """panel_out = (1/M) X diag(inv^2) X^T panel_in (src/native/gwldcore.cpp:1134-1257)."""
    _apply_grm("apply_grm_bed_panel", bed_prefix, fam_path, nsnps, step_size, row_sel, ddof, inv_all, panel_in,
               panel_out, C, R, impute_mode, impute_seed)


def apply_grm_bed_panel_mailman(bed_prefix, fam_path, nsnps, step_size, row_sel=None, ddof=1, *, inv_all, panel_in,
                                panel_out, C=None, R=None, impute_seed=None):
    """HWE-completed GRM apply (src/native/gwldcore.cpp:1259-1497): the same product
    as apply_grm_bed_panel with impute_mode="hwe" (the Mailman decomposition of the
    reference is replaced by the exact int8-MFMA products)."""
    _apply_grm("apply_grm_bed_panel_mailman", bed_prefix, fam_path, nsnps, step_size, row_sel, ddof, inv_all,
               panel_in, panel_out, C, R, "hwe", impute_seed)
