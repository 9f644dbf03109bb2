"""Device-level Python front end of libsummit_gpe.

``Genotype`` owns one genotype matrix resident in HBM (both packed layouts and
per-SNP statistics).  The functions below take torch tensors on the GPU and
hand their device pointers to the C ABI; PyTorch is used only to allocate and
move device buffers.  Shapes follow the reference's arrays:

* ``xz``: the reference's Fortran (N, Q) ``Xz2d`` is held as a contiguous
  torch tensor of shape (Q, N) (same memory layout).
* ``meansq``: (M, B) row-major FP64.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from . import _native
from ._native import c_i64, c_vp, call

IMPUTE = {"mean": 0, "hwe": 1}
RAND = {"rademacher": 0, "gaussian": 1, "normal": 1, "spherical": 2}


def rand_code(rand_dist: str, fp32: bool = False) -> int:
    """The reference's rule (src/native/gwldcore.cpp:1695-1696): "rademacher",
    "spherical", anything else Gaussian; float32 callers draw float variates."""
    if rand_dist == "rademacher":
        return 0
    if rand_dist == "spherical":
        return 4 if fp32 else 2
    return 3 if fp32 else 1
LAYOUT = {"snp_raw": 0, "snp_work": 1, "sample_work": 2}


def _ptr(t) -> int:
    if t is None:
        return 0
    return int(t.data_ptr())


def _ld(t) -> int:
    """Leading dimension of a row-major 2-D tensor; a single row's stride is
    arbitrary in torch (size-1 dimensions ignore it), so it is the row length."""
    return int(t.stride(0)) if t.shape[0] > 1 else max(int(t.stride(0)), int(t.shape[1]))


def _stream(stream=None) -> int:
    if stream is None:
        stream = torch.cuda.current_stream()
    return int(stream.cuda_stream)


def _np_ptr(a) -> int:
    return 0 if a is None else int(a.ctypes.data)


class Genotype:
    """A genotype matrix resident in HBM (n_samples x n_snps, 2-bit codes)."""

    def __init__(self, n_samples: int, n_snps: int):
        lib = _native.load()
        h = c_vp()
        _native.check(lib.gpe_geno_create(int(n_samples), int(n_snps), ctypes.byref(h)))
        self._h = h
        self.n = int(n_samples)
        self.m = int(n_snps)
        self.prepared_with = None

    # ------------------------------------------------------------------ creation
    @classmethod
    def from_bed_bytes(cls, body: np.ndarray, n_total: int, n_snps: int, rows=None, stream=None) -> "Genotype":
        """body: the BED bytes after the magic, (n_snps, ceil(n_total/4)) uint8 (may be a memmap)."""
        body = np.asarray(body)
        stride = (n_total + 3) // 4
        if body.ndim != 2 or body.shape[0] < n_snps or body.shape[1] != stride:
            raise RuntimeError("BED body must be (n_snps, ceil(n_total/4)) bytes")
        rows_arr = None
        n = n_total
        if rows is not None:
            rows_arr = np.ascontiguousarray(np.asarray(rows, dtype=np.int64))
            n = int(rows_arr.size)
        g = cls(n, n_snps)
        # a C-contiguous body (file mapping or array) goes down in one call, which
        # streams it through double-buffered pinned staging; other layouts are
        # made contiguous 256 MB at a time
        chunk = n_snps if body[:n_snps].flags.c_contiguous else max(1, (256 << 20) // max(stride, 1))
        for j0 in range(0, n_snps, chunk):
            j1 = min(n_snps, j0 + chunk)
            part = np.ascontiguousarray(body[j0:j1])
            call("gpe_geno_upload_bed", g._h, _np_ptr(part), n_total, stride, j0, j1 - j0, _np_ptr(rows_arr),
                 _stream(stream))
        return g

    @classmethod
    def synthetic(cls, n_samples: int, n_snps: int, seed: int, maf_lo: float = 0.05, maf_hi: float = 0.5,
                  miss_rate: float = 0.01, block_len: int = 100, copy_prob: float = 0.0, stream=None):
        g = cls(n_samples, n_snps)
        call("gpe_geno_synthesize", g._h, int(seed) & 0xFFFFFFFFFFFFFFFF, float(maf_lo), float(maf_hi),
             float(miss_rate), int(block_len), float(copy_prob), _stream(stream))
        return g

    # ------------------------------------------------------------------ state
    def prepare(self, ddof: int = 1, impute: str = "mean", impute_seed: int = 0, snp_offset: int = 0,
                global_rows=None, stream=None) -> "Genotype":
        if impute not in IMPUTE:
            raise RuntimeError("impute_mode must be 'mean' or 'hwe'")
        rows = None if global_rows is None else np.ascontiguousarray(np.asarray(global_rows, dtype=np.int64))
        call("gpe_geno_prepare", self._h, int(ddof), IMPUTE[impute], int(impute_seed) & 0xFFFFFFFFFFFFFFFF,
             int(snp_offset), _np_ptr(rows), _stream(stream))
        self.prepared_with = (int(ddof), impute, int(impute_seed))
        return self

    def stats(self) -> dict:
        nobs = np.empty(self.m, dtype=np.int32)
        s = np.empty(self.m, dtype=np.int64)
        ss = np.empty(self.m, dtype=np.int64)
        mean = np.empty(self.m, dtype=np.float64)
        inv_sd = np.empty(self.m, dtype=np.float64)
        call("gpe_geno_stats", self._h, _np_ptr(nobs), _np_ptr(s), _np_ptr(ss), _np_ptr(mean), _np_ptr(inv_sd))
        return {"nobs": nobs, "sum": s, "sumsq": ss, "mean": mean, "inv_sd": inv_sd}

    def n_missing(self) -> int:
        n, m, miss = c_i64(), c_i64(), c_i64()
        call("gpe_geno_shape", self._h, ctypes.byref(n), ctypes.byref(m), ctypes.byref(miss))
        return int(miss.value)

    def download(self, layout: str = "snp_raw") -> np.ndarray:
        if layout == "sample_work":
            rows, used = self.n, (self.m + 3) // 4
        else:
            rows, used = self.m, (self.n + 3) // 4
        out = np.empty((rows, used), dtype=np.uint8)
        call("gpe_geno_download", self._h, LAYOUT[layout], _np_ptr(out), used)
        return out

    def download_rows(self, layout: str, row_begin: int, row_count: int) -> np.ndarray:
        used = (self.m + 3) // 4 if layout == "sample_work" else (self.n + 3) // 4
        out = np.empty((row_count, used), dtype=np.uint8)
        call("gpe_geno_download_rows", self._h, LAYOUT[layout], int(row_begin), int(row_count), _np_ptr(out), used)
        return out

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _native.load().gpe_geno_free(self._h)
            self._h = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h


# ---------------------------------------------------------------------- products
def gz_i32(g: Genotype, z: torch.Tensor, snp_begin: int = 0, snp_end: int | None = None, out=None,
           accumulate: bool = False, stream=None) -> torch.Tensor:
    """U[b, n] = sum_i g[i, n] * z[i - snp_begin, b]; z int8 (L, B) on the GPU."""
    snp_end = g.m if snp_end is None else snp_end
    if z.dtype != torch.int8 or z.dim() != 2 or z.shape[0] != snp_end - snp_begin:
        raise RuntimeError("z must be int8 of shape (snp_end - snp_begin, B)")
    z = z.contiguous()
    B = z.shape[1]
    if out is None:
        out = torch.zeros((B, g.n), dtype=torch.int32, device=z.device)
        accumulate = True
    call("gpe_gz_i32", g.handle, snp_begin, snp_end, _ptr(z), B, B, _ptr(out), _ld(out), int(accumulate),
         _stream(stream))
    return out


def gtu_i64(g: Genotype, u: torch.Tensor, snp_begin: int = 0, snp_end: int | None = None,
            stream=None) -> torch.Tensor:
    """W[i - snp_begin, b] = sum_n g[i, n] * u[b, n] (exact int64)."""
    snp_end = g.m if snp_end is None else snp_end
    if u.dtype != torch.int32 or u.dim() != 2 or u.shape[1] != g.n:
        raise RuntimeError("u must be int32 of shape (B, n_samples)")
    u = u.contiguous()
    B = u.shape[0]
    w = torch.empty((snp_end - snp_begin, B), dtype=torch.int64, device=u.device)
    call("gpe_gtu_i64", g.handle, snp_begin, snp_end, _ptr(u), _ld(u), B, _ptr(w), B, _stream(stream))
    return w


def resvar_inv(g: Genotype, snp_begin: int = 0, snp_end: int | None = None, eps: float = 1e-10,
               out: torch.Tensor | None = None, stream=None) -> torch.Tensor:
    snp_end = g.m if snp_end is None else snp_end
    if out is None:
        out = torch.empty(snp_end - snp_begin, dtype=torch.float64, device="cuda")
    call("gpe_resvar_inv", g.handle, snp_begin, snp_end, float(eps), _ptr(out), _stream(stream))
    return out


def phase1_xz(g: Genotype, snp_begin: int, snp_end: int, step_size: int, annot: torch.Tensor,
              inv: torch.Tensor, v_start: int, v_count: int, seed: int, xz: torch.Tensor,
              rand_dist: str = "rademacher", bin_mode: torch.Tensor | None = None, limbs: int = 5,
              stream=None, probe_fp32: bool = False, probe_tile: int | None = None,
              v_start_dev: torch.Tensor | None = None) -> torch.Tensor:
    """Xz[k*v_count+v, :] (+)= sum_i x_i inv_i sqrt(a_ik) z_iv over [snp_begin, snp_end).
    probe_tile: probe columns come in V-tiles of this many, each with its own
    per-block streams (as separate calls of that width would draw)."""
    L = snp_end - snp_begin
    annot = annot.to(torch.float64).contiguous()
    inv = inv.to(torch.float64).contiguous()
    if annot.dim() != 2 or annot.shape[0] != L:
        raise RuntimeError("L mismatch")
    if inv.numel() != L:
        raise RuntimeError("inv_right shape mismatch in phase1")
    B = annot.shape[1]
    if xz.dtype != torch.float64 or xz.dim() != 2 or xz.shape[0] != B * v_count or xz.shape[1] != g.n:
        raise RuntimeError("Xz2d_chunk shape must be (N, B*v_count)")
    tile = v_count if probe_tile is None else int(probe_tile)
    if v_start_dev is not None:
        # first probe id read on the device (graph-capturable; Rademacher only)
        if rand_dist != "rademacher" or v_start_dev.dtype != torch.int32 or v_start_dev.numel() != 1:
            raise RuntimeError("v_start_dev needs Rademacher probes and a one-element int32 device tensor")
        call("gpe_phase1_xz_indirect", g.handle, snp_begin, snp_end, step_size, _ptr(annot), B, _ptr(inv),
             _ptr(v_start_dev), v_count, tile, int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(xz), _ld(xz),
             _ptr(bin_mode), limbs, _stream(stream))
        return xz
    call("gpe_phase1_xz_tiled", g.handle, snp_begin, snp_end, step_size, _ptr(annot), B, _ptr(inv), v_start, v_count,
         tile, rand_code(rand_dist, probe_fp32), int(seed) & 0xFFFFFFFFFFFFFFFF, _ptr(xz), _ld(xz),
         _ptr(bin_mode), limbs, _stream(stream))
    return xz


def add_i32(value: torch.Tensor, increment: int, stream=None) -> None:
    """value[0] += increment on the device, in stream order (graph-capturable)."""
    if value.dtype != torch.int32 or value.numel() < 1:
        raise RuntimeError("value must be an int32 device tensor")
    call("gpe_add_i32", _ptr(value), int(increment), _stream(stream))


def phase2_accum(g: Genotype, snp_begin: int, snp_end: int, inv_left: torch.Tensor, tile_nvecs: int,
                 xz: torch.Tensor, meansq: torch.Tensor, n_denom: int, mc_aggregate_sumsq=None,
                 mc_per_snp_sumsq=None, extreme_warn: float = 1.0e10, drop_extreme: bool = True,
                 limbs: int = 5, meansq_row0: int = 0, stream=None, want_guards: bool = False, sum_xz=None):
    """meansq[j, k] += sum_v (x_j . Xz[k*V+v])^2 inv_j^2/(N_denom-1)^2 for j in [snp_begin, snp_end).

    ``meansq`` may hold only rows [meansq_row0, ...); it is addressed by absolute SNP index.
    ``sum_xz`` (device FP64, one value per column of Xz) replaces the column sums of
    Xz in the centring x_j . Xz_c = inv_sd_j (g_j . Xz_c - mu_j sum_xz[c])."""
    if xz.dtype != torch.float64 or xz.dim() != 2 or xz.shape[1] != g.n:
        raise RuntimeError("Xz2d shape mismatch in phase2_accum_XtXz_bed")
    Q = xz.shape[0]
    if tile_nvecs <= 0 or Q % tile_nvecs:
        raise RuntimeError("Xz2d col count must be a multiple of tile_nvecs")
    B = Q // tile_nvecs
    if meansq.dtype != torch.float64 or meansq.dim() != 2 or meansq.shape[1] != B or not meansq.is_contiguous():
        raise RuntimeError("meansq_accum shape mismatch in phase2_accum_XtXz_bed")
    inv_left = inv_left.to(torch.float64).contiguous()
    if inv_left.numel() != snp_end - snp_begin:
        raise RuntimeError("inv_left shape mismatch in phase2_accum_XtXz_bed")
    shift = meansq_row0 * B * 8
    base = _ptr(meansq) - shift
    mcp = 0 if mc_per_snp_sumsq is None else _ptr(mc_per_snp_sumsq) - shift
    counts = (ctypes.c_ulonglong * 2)()
    guards = ctypes.cast(counts, c_vp) if want_guards else None
    if sum_xz is None:
        call("gpe_phase2_accum", g.handle, snp_begin, snp_end, _ptr(inv_left), tile_nvecs, _ptr(xz), _ld(xz),
             B, int(n_denom), base, B, _ptr(mc_aggregate_sumsq), mcp, float(extreme_warn), int(bool(drop_extreme)),
             limbs, guards, _stream(stream))
    else:
        sum_xz = sum_xz.to(torch.float64).contiguous()
        if sum_xz.numel() != Q:
            raise RuntimeError("sum_Xz shape mismatch in phase2_accum_XtXz_bed_mailman_rm")
        call("gpe_phase2_accum_colsums", g.handle, snp_begin, snp_end, _ptr(inv_left), tile_nvecs, _ptr(xz),
             _ld(xz), B, _ptr(sum_xz), int(n_denom), base, B, _ptr(mc_aggregate_sumsq), mcp,
             float(extreme_warn), int(bool(drop_extreme)), limbs, guards, _stream(stream))
    if want_guards:
        return int(counts[0]), int(counts[1])
    return None


def xc(g: Genotype, c: torch.Tensor, snp_begin: int = 0, snp_end: int | None = None, out=None,
       accumulate: bool = False, limbs: int = 5, stream=None) -> torch.Tensor:
    """out[q, n] (+)= sum_i x[n, i] c[i - snp_begin, q] (standardized X; out (Q, N) = column-major (N, Q))."""
    snp_end = g.m if snp_end is None else snp_end
    c = c.to(torch.float64).contiguous()
    if c.dim() != 2 or c.shape[0] != snp_end - snp_begin:
        raise RuntimeError("c must be (snp_end - snp_begin, Q)")
    Q = c.shape[1]
    if out is None:
        out = torch.zeros((Q, g.n), dtype=torch.float64, device=c.device)
    call("gpe_xc_f64", g.handle, snp_begin, snp_end, _ptr(c), Q, Q, _ptr(out), _ld(out), int(bool(accumulate)),
         limbs, _stream(stream))
    return out


def xtv(g: Genotype, v: torch.Tensor, snp_begin: int = 0, snp_end: int | None = None, limbs: int = 5,
        stream=None) -> torch.Tensor:
    """out[j - snp_begin, q] = sum_n x[n, j] v[q, n]; v (Q, N) = column-major (N, Q)."""
    snp_end = g.m if snp_end is None else snp_end
    v = v.to(torch.float64).contiguous()
    if v.dim() != 2 or v.shape[1] != g.n:
        raise RuntimeError("v must be (Q, n_samples)")
    Q = v.shape[0]
    out = torch.empty((snp_end - snp_begin, Q), dtype=torch.float64, device=v.device)
    call("gpe_xtv_f64", g.handle, snp_begin, snp_end, _ptr(v), _ld(v), Q, _ptr(out), Q, limbs,
         _stream(stream))
    return out


def sq_rowsums(g: Genotype, w_kl: torch.Tensor, snp_begin: int = 0, snp_end: int | None = None,
               C: torch.Tensor | None = None, rx: torch.Tensor | None = None, out=None, stream=None) -> torch.Tensor:
    """out[k, i] += sum_j xt[i, j]^2 w_kl[k, j - snp_begin] over SNPs j of [snp_begin, snp_end): per-sample
    sums of squared standardized genotypes, projected xt = x - C (R x) when C (N, p) and rx (L, p) = (R x_j)_a
    are given.  w_kl (K, L) FP64 = column-major (L, K); out (K, N) FP64, zeros when None."""
    snp_end = g.m if snp_end is None else snp_end
    L = snp_end - snp_begin
    w_kl = w_kl.to(torch.float64).contiguous()
    if w_kl.dim() != 2 or w_kl.shape[1] != L:
        raise RuntimeError("weights must be (K, snp_end - snp_begin)")
    K = w_kl.shape[0]
    if out is None:
        out = torch.zeros((K, g.n), dtype=torch.float64, device=w_kl.device)
    if out.shape != (K, g.n) or out.dtype != torch.float64 or out.stride(1) != 1:
        raise RuntimeError("out must be (K, n_samples) FP64")
    p = 0
    c_ptr = r_ptr = 0
    ldc = ldr = 0
    if C is not None:
        c_cm = C.to(torch.float64).t().contiguous()  # (p, N) rows = columns of C
        rx = rx.to(torch.float64).contiguous()
        p = int(c_cm.shape[0])
        if c_cm.shape[1] != g.n or rx.shape != (L, p):
            raise RuntimeError("C must be (N, p) and rx (snp_end - snp_begin, p)")
        c_ptr, ldc, r_ptr, ldr = _ptr(c_cm), g.n, _ptr(rx), p
    call("gpe_sq_rowsums_f64", g.handle, snp_begin, snp_end, _ptr(w_kl), _ld(w_kl), K, c_ptr, ldc, r_ptr, ldr, p,
         _ptr(out), _ld(out), _stream(stream))
    return out


def project(x_qn: torch.Tensor, A: torch.Tensor | None, B: torch.Tensor | None, center: bool = False,
            stream=None) -> torch.Tensor:
    """In place: x_q <- x_q - A (B x_q) [then minus its mean] for the rows q of
    x_qn (Q, N) FP64 (= column-major (N, Q)); A (N, p), B (p, N) FP64 device
    tensors or None (centring only).  Returns x_qn."""
    if x_qn.dtype != torch.float64 or x_qn.dim() != 2 or x_qn.stride(1) != 1:
        raise RuntimeError("projected matrix must be (Q, N) FP64 with unit stride along N")
    Q, N = x_qn.shape
    p = 0
    a_ptr = b_ptr = 0
    lda = ldb = N
    if A is not None:
        A_cm = A.to(torch.float64).t().contiguous()  # (p, N) rows = columns of A
        B_rm = B.to(torch.float64).contiguous()
        if A_cm.shape[1] != N or B_rm.shape != A_cm.shape:
            raise RuntimeError("projection factors must be (N, p) and (p, N)")
        p = int(A_cm.shape[0])
        a_ptr, b_ptr = _ptr(A_cm), _ptr(B_rm)
    if p == 0 and not center:
        return x_qn
    call("gpe_project_f64", _ptr(x_qn), _ld(x_qn), N, Q, a_ptr, lda, b_ptr, ldb, p, int(bool(center)),
         _stream(stream))
    if A is not None and stream is not None:
        # the factor copies were made on torch's current stream; keep them until `stream` is done
        A_cm.record_stream(stream)
        B_rm.record_stream(stream)
    return x_qn


def col_sums(x_qn: torch.Tensor, stream=None) -> torch.Tensor:
    """sum over N of each row q of x_qn (Q, N) FP64 (= column sums of column-major (N, Q))."""
    if x_qn.dtype != torch.float64 or x_qn.dim() != 2 or x_qn.stride(1) != 1:
        raise RuntimeError("column sums need a (Q, N) FP64 matrix with unit stride along N")
    Q, N = x_qn.shape
    out = torch.empty(Q, dtype=torch.float64, device=x_qn.device)
    call("gpe_col_sums_f64", _ptr(x_qn), _ld(x_qn), N, Q, _ptr(out), _stream(stream))
    return out


def basis_coef(x_qn: torch.Tensor, b_pn: torch.Tensor, stream=None) -> tuple[torch.Tensor, torch.Tensor]:
    """(t, ss): t[k, q] = sum_n b_pn[k, n] x_qn[q, n] and ss[q] = sum_n x_qn[q, n]^2 for x_qn (Q, N)
    and b_pn (p, N) FP64 with unit stride along N (fixed-order reductions)."""
    for t in (x_qn, b_pn):
        if t.dtype != torch.float64 or t.dim() != 2 or t.stride(1) != 1:
            raise RuntimeError("basis coefficients need (rows, N) FP64 matrices with unit stride along N")
    Q, N = x_qn.shape
    p = b_pn.shape[0]
    if b_pn.shape[1] != N:
        raise RuntimeError("basis and panel disagree in N")
    t_out = torch.empty((max(p, 1), Q), dtype=torch.float64, device=x_qn.device)
    ss = torch.empty(Q, dtype=torch.float64, device=x_qn.device)
    call("gpe_basis_coef_f64", _ptr(x_qn), _ld(x_qn), N, Q, _ptr(b_pn), _ld(b_pn), p, _ptr(t_out), Q, _ptr(ss),
         _stream(stream))
    return t_out[:p], ss


def apply_grm(g: Genotype, panel: torch.Tensor, inv_all: torch.Tensor, nsnps: int | None = None,
              step_size: int | None = None, limbs: int = 5, stream=None) -> torch.Tensor:
    """panel_out = (1/M) X diag(inv^2) X^T panel_in over the first nsnps SNPs
    (apply_grm_bed_panel, src/native/gwldcore.cpp:1134-1257).  panel (Q, N) = column-major (N, Q).
    The whole SNP range is one product pair; step_size only bounds the coefficient panel."""
    M = g.m if nsnps is None else int(nsnps)
    step = M if step_size is None else max(1, int(step_size))
    inv_all = inv_all.to(torch.float64)
    w = inv_all[:M] * inv_all[:M] / float(M)
    w = torch.where(torch.isfinite(w) & (w > 0), w, torch.zeros_like(w))
    out = torch.zeros_like(panel, dtype=torch.float64)
    for s in range(0, M, step):
        e = min(M, s + step)
        coef = xtv(g, panel, s, e, limbs=limbs, stream=stream)
        coef *= w[s:e, None]
        xc(g, coef, s, e, out=out, accumulate=True, limbs=limbs, stream=stream)
    return out


def probes(root_seed: int, blk_starts, blk_end: int, v_start: int, v_count: int,
           rand_dist: str = "rademacher", fp32: bool = False) -> np.ndarray:
    """Reference probe values for consecutive blocks, (blk_end - blk_starts[0], v_count) FP64."""
    starts = np.ascontiguousarray(np.asarray(blk_starts, dtype=np.int64))
    out = np.empty((blk_end - int(starts[0]), v_count), dtype=np.float64)
    call("gpe_probes", rand_code(rand_dist, fp32), int(root_seed) & 0xFFFFFFFFFFFFFFFF, _np_ptr(starts), len(starts),
         int(blk_end), int(v_start), int(v_count), _np_ptr(out))
    return out


PROF_KINDS = {"p1_contract": 0, "p2_contract": 1, "gz": 2, "gtu": 3, "gram": 4, "p1_finalize": 5, "p2_miss": 6}


def profile_enable(on: bool = True) -> None:
    call("gpe_profile_enable", int(bool(on)))


def profile_reset() -> None:
    call("gpe_profile_reset")


def profile_read(kind: str) -> tuple[float, int]:
    """(total milliseconds, launches) of one kernel kind since the last reset."""
    ms, n = ctypes.c_double(), c_i64()
    call("gpe_profile_read", PROF_KINDS[kind], ctypes.byref(ms), ctypes.byref(n))
    return float(ms.value), int(n.value)


def default_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("summit_amd needs a HIP device (torch.cuda.is_available() is False)")
    return torch.device("cuda", torch.cuda.current_device())


os.environ.setdefault("HSA_ENABLE_IPC_MODE_LEGACY", "0")
