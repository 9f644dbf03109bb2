"""Drop-in mirror of the reference's ``gxeldcore.DirectContext`` / ``ProjectedPanel``.

The per-block GxE backend that GenomewideEnvLDScore drives
(src/summit/ldscore/gwe_ldscore.py:1406-1440, :2218-2260, :2716-2830), with the
constructor, method names, keyword arguments, return values and error texts of
src/native/gxeldcore.cpp:7086-8877 (bindings :16541-16593).  The work runs in
libsummit_gpe on the GPU:

* construction makes the selected samples of the BED body behind the descriptor
  resident in HBM once; every block is served from there (the reference
  memory-maps the file and decodes each block on the host);
* features depend on the SNP alone, so the first call that needs them computes
  every SNP's: the per-SNP sums sum_n e_n^p x_nj^2 (p = 0, 1, 2, 4) in one pass
  (gpe_weighted_sumsq) and the moments [Q, EQ, E^2 Q, E^3 Q]^T X in one
  phase-2 traversal (gpe_xtv_f64), the per-SNP algebra on them in FP64
  (gpe_gxe_direct_features, one thread per SNP); each
  block's call returns its slice, checked as the reference checks it;
* sources X diag(a s) Z: one phase-1 traversal of FP64 coefficients (gpe_xc_f64,
  six int8 limbs), then e . and the projection I - Q Q^T (gpe_project_f64),
  returned in pinned host memory (one DMA);
* targets X^T [R, e . R]: one phase-2 traversal over both halves (gpe_xtv_f64);
* a ProjectedPanel keeps its [S, e . S] in HBM, so target blocks re-read it there.

The engine standardizes x_eng = (g - mean) inv_eng with the variance over the
observed calls (nobs - ddof); DirectContext standardizes the A1 dosage over the
selected samples (N - ddof): x = -r x_eng with r = inv_dc / inv_eng per SNP
(1 without missing calls).  Products are taken on x_eng and scaled by -r.

Differences a caller can see:

* ``decode_threads`` / ``blas_threads`` are validated and reported, not used;
* the duplicated feature-moment product of ``strict_feature_moment_verification``
  is not run: the moments are exact integer contractions, deterministic and
  repeatable, so a duplicate cannot disagree; the repair counters stay 0;
* workspace limits use the reference's element counts per call without the CPU
  GEMM-integrity term (0 in reference builds without that check);
* results carry the engine's fixed-point coefficient rounding (six limbs,
  relative 2^-44 of each coefficient column's largest entry) instead of BLAS
  rounding.
"""
from __future__ import annotations

import fcntl
import itertools
import math
import mmap
import os
import stat
import threading

import numpy as np
import torch

from . import engine
from .gxe import weighted_sumsq

LIMBS = 6
_ENV_TOL = 1.0e-8  # kStandardizedEnvTolerance (gxeldcore.cpp:131)
_ORTHO_TOL = 1.0e-10  # kOrthonormalTolerance (gxeldcore.cpp:130)
_INT_MAX = 2**31 - 1
_DBL_MIN = float(np.finfo(np.float64).tiny)
_IDS = itertools.count(1)
_WHITESPACE = np.array([32, 9, 13, 12, 11], dtype=np.uint8)


# --------------------------------------------------------------------------- descriptors
def _dup_cloexec(fd, label: str) -> int:
    """duplicate_cloexec (gxeldcore.cpp:6460-6485)."""
    fd = int(fd)
    if fd < 0:
        raise RuntimeError(f"Invalid GxE native {label} descriptor")
    try:
        return fcntl.fcntl(fd, fcntl.F_DUPFD_CLOEXEC, 0)
    except OSError as exc:
        raise RuntimeError(f"Failed to duplicate GxE native {label} descriptor with close-on-exec: "
                           f"{os.strerror(exc.errno)}") from None


def _fd_state(fd: int, label: str) -> tuple:
    """validate_regular_fd / state_from_stat (gxeldcore.cpp:6438-6501)."""
    try:
        st = os.fstat(fd)
    except OSError as exc:
        raise RuntimeError(f"Failed to stat GxE native {label} descriptor: {os.strerror(exc.errno)}") from None
    if not stat.S_ISREG(st.st_mode) or st.st_size < 0:
        raise RuntimeError(f"GxE native {label} descriptor must reference a regular file")
    return (st.st_dev, st.st_ino, st.st_size, st.st_nlink, st.st_mtime_ns, st.st_ctime_ns)


def _count_rows(fd: int, size: int, label: str) -> int:
    """count_validated_rows_fd (gxeldcore.cpp:6503-6553): newline-terminated rows,
    each holding a byte other than blank, tab, CR, FF or VT; a started final row
    without a newline counts when it has content."""
    data = bytearray()
    off = 0
    while off < size:
        chunk = os.pread(fd, min(1 << 20, size - off), off)
        if not chunk:
            raise RuntimeError(f"Unexpected EOF in GxE native {label}")
        data += chunk
        off += len(chunk)
    buf = np.frombuffer(bytes(data), dtype=np.uint8)
    if buf.size == 0:
        raise RuntimeError(f"GxE native {label} has invalid row count")
    newline = buf == 10
    content = ~newline & ~np.isin(buf, _WHITESPACE)
    line = np.cumsum(newline) - newline  # row index of each byte (a newline belongs to its row)
    nl = int(newline.sum())
    has = np.bincount(line[content], minlength=nl + 1) > 0
    if nl and not bool(has[:nl].all()):
        raise RuntimeError(f"GxE native {label} contains a blank row")
    rows = nl
    if buf[-1] != 10:  # a final row was started
        if not bool(has[nl]):
            raise RuntimeError(f"GxE native {label} contains a blank final row")
        rows += 1
    if rows <= 0 or rows > _INT_MAX:
        raise RuntimeError(f"GxE native {label} has invalid row count")
    return rows


def _parse_rows(row_sel, n_total: int) -> np.ndarray:
    """DirectContext::parse_rows (gxeldcore.cpp:8504-8546): a snapshot of the selection."""
    if row_sel is None:
        rows = np.arange(n_total, dtype=np.int64)
    else:
        idx = np.asarray(row_sel)
        if idx.ndim != 1 or idx.shape[0] > _INT_MAX:
            raise RuntimeError("GxE native row_sel must be a bounded one-dimensional integer array")
        if idx.dtype == np.int32:
            rows = idx.astype(np.int64)
        elif idx.dtype == np.int64:
            if idx.size and (int(idx.min()) < 0 or int(idx.max()) > _INT_MAX):
                raise RuntimeError("GxE native row_sel contains an out-of-range sample index")
            rows = idx.copy()
        else:
            raise RuntimeError("GxE native row_sel must have dtype int32 or int64")
    if rows.size < 3:
        raise RuntimeError("GxE native context requires at least three selected samples")
    if rows.size and (int(rows[0]) < 0 or int(rows[-1]) >= n_total or bool(np.any(np.diff(rows) <= 0))):
        raise RuntimeError("GxE native row_sel must be strictly increasing and within the FAM axis")
    return np.ascontiguousarray(rows)


# --------------------------------------------------------------------------- panel
class ProjectedPanel:
    """An opaque, immutable [S, e . S] panel bound to the DirectContext that sealed
    it (gxeldcore.cpp:6599-6633); its data stays on the GPU."""

    __slots__ = ("_ctx", "_rows", "_columns", "_leakage", "_data")

    def __init__(self, *args, **kwargs):
        raise TypeError("gxeldcore.ProjectedPanel: no constructor defined!")

    @classmethod
    def _seal(cls, ctx: int, rows: int, columns: int, leakage: float, data: torch.Tensor) -> "ProjectedPanel":
        panel = object.__new__(cls)
        panel._ctx, panel._rows, panel._columns, panel._leakage, panel._data = ctx, rows, columns, leakage, data
        return panel

    @property
    def columns(self) -> int:
        return self._columns

    @property
    def leakage(self) -> float:
        return self._leakage


# --------------------------------------------------------------------------- context
class DirectContext:
    """gxeldcore.DirectContext (gxeldcore.cpp:7086-8877) on the GPU engine."""

    def __init__(self, bed_descriptor, bim_descriptor, fam_descriptor, row_sel=None, ddof=1, env=None,
                 q_basis=None, decode_threads=None, max_workspace_bytes=None, target_panel_columns=64,
                 strict_feature_moment_verification=True, blas_threads=0):
        if env is None or q_basis is None or decode_threads is None or max_workspace_bytes is None:
            raise TypeError("DirectContext() needs env, q_basis, decode_threads and max_workspace_bytes")
        self._closed = True
        self._fds: list[int] = []
        self._map = None
        self._g = None
        self._ftab = None
        self._lock = threading.Lock()
        self._id = next(_IDS)
        self._ddof = int(ddof)
        self._decode_threads = int(decode_threads)
        self._blas_threads = int(blas_threads) if int(blas_threads) > 0 else int(decode_threads)
        if int(max_workspace_bytes) < 0:
            raise TypeError("max_workspace_bytes must be a non-negative integer")
        self._max_ws = int(max_workspace_bytes)
        self._tpc = int(target_panel_columns)
        self._strict = bool(strict_feature_moment_verification)
        if self._ddof not in (0, 1):
            raise RuntimeError("GxE native context supports ddof 0 or 1")
        if self._decode_threads <= 0:
            raise RuntimeError("GxE native decode_threads must be positive")
        if self._blas_threads <= 0:
            raise RuntimeError("GxE native blas_threads must be positive")
        limit = len(os.sched_getaffinity(0))
        if self._decode_threads > limit:
            raise RuntimeError("GxE native decode_threads exceeds the OpenMP/CPU-affinity limit")
        if self._blas_threads > limit:
            raise RuntimeError("GxE native blas_threads exceeds the OpenMP/CPU-affinity limit")
        if self._max_ws == 0 or self._max_ws > 2**64 - 1:
            raise RuntimeError("GxE native max_workspace_bytes is invalid")
        if self._tpc <= 0:
            raise RuntimeError("GxE native target_panel_columns must be positive")
        self._closed = False
        try:
            self._open(bed_descriptor, bim_descriptor, fam_descriptor, row_sel, env, q_basis)
        except BaseException:
            self._close_internal()
            raise

    # ------------------------------------------------------------------ construction
    def _open(self, bed_fd, bim_fd, fam_fd, row_sel, env, q_basis) -> None:
        for fd, label in ((bed_fd, "BED"), (bim_fd, "BIM"), (fam_fd, "FAM")):
            self._fds.append(_dup_cloexec(fd, label))
        labels = ("BED", "BIM", "FAM")
        self._states = [_fd_state(fd, lab) for fd, lab in zip(self._fds, labels)]
        self._n_total = _count_rows(self._fds[2], self._states[2][2], "FAM")
        self._m_total = _count_rows(self._fds[1], self._states[1][2], "BIM")
        stride = (self._n_total + 3) // 4
        expected = 3 + stride * self._m_total
        if self._states[0][2] != expected:
            raise RuntimeError("GxE native BED byte size disagrees with FAM/BIM dimensions")
        try:
            self._map = mmap.mmap(self._fds[0], expected, flags=mmap.MAP_PRIVATE, prot=mmap.PROT_READ)
        except (OSError, ValueError) as exc:
            raise RuntimeError(f"Failed to mmap GxE native BED: {exc}") from None
        if self._map[:3] != b"\x6c\x1b\x01":
            raise RuntimeError("GxE native input is not a SNP-major PLINK BED")
        self._rows = _parse_rows(row_sel, self._n_total)
        self._n = int(self._rows.size)
        self._design(env, q_basis)
        self._compare_states()
        self._upload(stride)

    def _design(self, env, q_basis) -> None:
        """copy_and_validate_design (gxeldcore.cpp:8547-8655): environment
        centred and standardized for ddof, Q orthonormal and spanning the intercept
        and the environment; the Q Gram and Q^T E^2 Q the feature algebra uses."""
        n = self._n
        env = np.array(env, dtype=np.float64).reshape(-1)
        q = np.array(q_basis, dtype=np.float64)
        if q.ndim != 2 or env.shape[0] != n or q.shape[0] != n:
            raise RuntimeError("GxE native environment/projection row mismatch")
        p = int(q.shape[1])
        if p < 1 or p >= n:
            raise RuntimeError("GxE native projection basis has invalid rank")
        if not bool(np.isfinite(env).all()):
            raise RuntimeError("GxE native environment contains a non-finite value")
        env_ss = math.fsum((env * env).tolist())
        self._env_mean = math.fsum(env.tolist()) / n
        self._env_var = env_ss / (n - self._ddof) if n - self._ddof else float("inf")
        if abs(self._env_mean) > _ENV_TOL or not math.isfinite(self._env_var) or abs(self._env_var - 1.0) > _ENV_TOL:
            raise RuntimeError("GxE native environment must be nonconstant, centered, and standardized for the "
                               "configured ddof")
        if not bool(np.isfinite(q).all()):
            raise RuntimeError("GxE native projection basis contains a non-finite value")
        gram = q.T @ q
        e2 = env * env
        e2gram = q.T @ (e2[:, None] * q)
        self._q_gram_err = float(np.max(np.abs(gram - np.eye(p))))
        if self._q_gram_err > _ORTHO_TOL:
            raise RuntimeError("GxE native projection basis is not orthonormal")
        root_n = math.sqrt(float(n))
        ri = 1.0 / root_n - q @ (q.T @ np.full(n, 1.0 / root_n))
        re = env - q @ (q.T @ env)
        if math.sqrt(float(ri @ ri)) > 1e-9 or math.sqrt(float(re @ re) / max(env_ss, _DBL_MIN)) > 1e-9:
            raise RuntimeError("GxE native projection basis must span the intercept and environment")
        self._q = p
        self._env = env
        dev = engine.default_device()
        self._dev = dev
        self._env_d = torch.as_tensor(env, device=dev)
        self._qt_d = torch.as_tensor(np.ascontiguousarray(q.T), device=dev)  # (p, N): Q column-major
        # feature moment basis rows [Q, E Q, E^2 Q, E^3 Q] (gxeldcore.cpp moment basis) and the
        # weights 1, e, e^2, e^4 of the per-SNP sums of squares
        mult = np.ones(n)
        basis = []
        for _ in range(4):
            basis.append(q.T * mult[None, :])
            mult = mult * env
        self._v4_d = torch.as_tensor(np.ascontiguousarray(np.vstack(basis)), device=dev)
        self._w4_d = torch.as_tensor(np.ascontiguousarray(np.vstack([np.ones(n), env, e2, e2 * e2])), device=dev)
        self._gram_d = torch.as_tensor(gram, device=dev)
        self._e2gram_d = torch.as_tensor(e2gram, device=dev)

    def _upload(self, stride: int) -> None:
        body = np.frombuffer(self._map, dtype=np.uint8, count=stride * self._m_total, offset=3)
        body = body.reshape(self._m_total, stride)
        self._g = engine.Genotype.from_bed_bytes(body, self._n_total, self._m_total, rows=self._rows)
        del body
        # every block is served from HBM from here on; the descriptors stay open for the
        # file-identity checks
        self._map.close()
        self._map = None
        self._g.prepare(ddof=self._ddof, impute="mean")
        st = self._g.stats()
        nobs = st["nobs"].astype(np.int64)
        sm = st["sum"].astype(np.float64)
        ss = st["sumsq"].astype(np.float64)
        dn = np.maximum(nobs, 1).astype(np.float64)
        m2 = np.where(nobs > 0, ss - sm * sm / dn, 0.0)
        m2 = np.where((m2 < 0.0) & (m2 > -1e-12), 0.0, m2)
        denom = self._n - self._ddof
        inv_dc = np.ones(self._m_total)
        ok = (denom > 0) & (m2 > 0.0)
        inv_dc[ok] = np.sqrt(float(denom) / m2[ok])
        # x = -r x_eng; where every observed call is equal both forms are 0 and r = 1
        r = np.where(m2 > 0.0, inv_dc / st["inv_sd"], 1.0)
        self._neg_r = -r
        self._neg_r_d = torch.as_tensor(self._neg_r, device=self._dev)
        miss = (self._n - nobs).astype(np.int64)
        self._miss_cum = np.concatenate([[0], np.cumsum(miss)])

    # ------------------------------------------------------------------ checks
    def _ensure_open(self) -> None:
        if self._closed or self._g is None:
            raise RuntimeError("GxE native context is closed")

    def _validate_block(self, blk_start: int, blk_end: int) -> None:
        self._ensure_open()
        if blk_start < 0 or blk_end <= blk_start or blk_end > self._m_total:
            raise RuntimeError("GxE native BED block is outside the BIM axis")

    def _check_files(self) -> None:
        self._ensure_open()
        self._compare_states()

    def _compare_states(self) -> None:
        """check_files_unchanged (gxeldcore.cpp:8812-8822): same device, inode, size,
        link count, mtime and ctime as at construction."""
        for fd, label, state in zip(self._fds, ("BED", "BIM", "FAM"), self._states):
            if _fd_state(fd, label) != state:
                raise RuntimeError("GxE native PLINK input changed after context construction")

    def _ensure_workspace(self, elements: int, label: str) -> None:
        need = elements * 8
        if need > self._max_ws:
            raise RuntimeError(f"GxE native {label} requires {need} workspace bytes, exceeding the configured limit "
                               f"{self._max_ws}")

    def _missing(self, blk_start: int, blk_end: int, require_missing_free: bool) -> int:
        missing = int(self._miss_cum[blk_end] - self._miss_cum[blk_start])
        if require_missing_free and missing != 0:
            raise RuntimeError("GxE native production backend requires a missing-free selected BED block")
        return missing

    def _panel_handle(self, panel) -> None:
        if not isinstance(panel, ProjectedPanel):
            raise TypeError("expected a gxeldcore.ProjectedPanel")
        if panel._ctx != self._id or panel._rows != self._n or panel._columns <= 0 or panel._data is None:
            raise RuntimeError("GxE native projected panel does not belong to this context")

    @staticmethod
    def _vec(a) -> np.ndarray:
        """A snapshot of a one-dimensional FP64 argument."""
        v = np.array(a, dtype=np.float64)
        if v.ndim != 1:
            raise RuntimeError("expected a one-dimensional array")
        return v

    @staticmethod
    def _validate_scales(sx: np.ndarray, sw: np.ndarray) -> None:
        if not bool((np.isfinite(sx) & np.isfinite(sw) & (sx > 0.0) & (sw > 0.0)).all()):
            raise RuntimeError("GxE native target feature scale is invalid")

    # ------------------------------------------------------------------ device helpers
    def _project(self, x_qn: torch.Tensor) -> torch.Tensor:
        """x_q <- x_q - Q (Q^T x_q) in place (project_panel_inplace, gxeldcore.cpp:8768-8781)."""
        qt = self._qt_d
        engine.call("gpe_project_f64", engine._ptr(x_qn), engine._ld(x_qn), self._n, int(x_qn.shape[0]),
                    engine._ptr(qt), self._n, engine._ptr(qt), self._n, self._q, 0, engine._stream())
        return x_qn

    def _upload_panel(self, sources) -> torch.Tensor:
        """(N, C) caller array -> (C, N) device rows (column-major copy)."""
        s = np.asarray(sources, dtype=np.float64)
        return torch.as_tensor(np.ascontiguousarray(s.T), device=self._dev)

    def _leakage(self, s_qn: torch.Tensor, width: int | None = None) -> float:
        """||Q^T S||_F / ||S||_F, the largest over panels of `width` columns when given
        (projected_source_leakage_from_coefficients, gxeldcore.cpp:8244-8287)."""
        t, ss = engine.basis_coef(s_qn, self._qt_d)
        t2 = (t * t).sum(dim=0).cpu().numpy()
        ssh = ss.cpu().numpy()
        C = t2.size
        width = C if width is None else width
        worst = 0.0
        for c0 in range(0, C, width):
            num = math.fsum(t2[c0:c0 + width].tolist())
            den = max(math.fsum(ssh[c0:c0 + width].tolist()), _DBL_MIN)
            leak = math.sqrt(num / den)
            if not math.isfinite(leak):
                raise RuntimeError("GxE native projected source diagnostic is non-finite")
            worst = max(worst, leak)
        return worst

    def _products(self, blk_start: int, blk_end: int, rows_qn: torch.Tensor) -> torch.Tensor:
        """x_j^T v_q for the block's SNPs j, the DirectContext's x: (L, Q) on the GPU."""
        out = engine.xtv(self._g, rows_qn, blk_start, blk_end, limbs=LIMBS)
        out.mul_(self._neg_r_d[blk_start:blk_end, None])
        return out

    _FEATURE_NAMES = ("scale_x", "scale_w", "norm_x", "norm_w", "diag_nxe_x", "diag_nxe_w", "corr_xw")

    def _feature_table(self) -> np.ndarray:
        """compute_feature_moments_from_genotype (gxeldcore.cpp:8033-8243) for every SNP
        at once, at the first call that needs features: one weighted-sum-of-squares
        pass and one moment traversal over the resident matrix serve every block.
        Rows: the seven feature vectors, varx, varw and the two leakage ratios (M,)."""
        if self._ftab is not None:
            return self._ftab
        M, p = self._m_total, self._q
        r2 = self._neg_r_d ** 2
        sums = weighted_sumsq(self._g, self._w4_d, 0, M)
        U = self._products(0, M, self._v4_d)
        tab = torch.empty((11, M), dtype=torch.float64, device=self._dev)
        engine.call("gpe_gxe_direct_features", engine._ptr(sums), engine._ld(sums), engine._ptr(U), engine._ld(U), p,
                    engine._ptr(self._gram_d), engine._ptr(self._e2gram_d), engine._ptr(r2), float(self._n - p), M,
                    engine._ptr(tab), M, engine._stream())
        self._ftab = tab.cpu().numpy()
        return self._ftab

    def _features(self, blk_start: int, blk_end: int, eps_var: float) -> dict:
        """The block's features (host arrays), with the reference's per-column checks in
        column order: zero or invalid variance, then non-finite outputs."""
        tab = self._feature_table()[:, blk_start:blk_end]
        k = len(self._FEATURE_NAMES)
        vx, vw = tab[k], tab[k + 1]
        bad_var = ~(np.isfinite(vx) & np.isfinite(vw) & (vx > eps_var) & (vw > eps_var))
        bad_out = ~np.isfinite(tab[:k]).all(axis=0)
        first_var = int(np.argmax(bad_var)) if bad_var.any() else None
        first_out = int(np.argmax(bad_out)) if bad_out.any() else None
        if first_var is not None and (first_out is None or first_var <= first_out):
            raise RuntimeError(f"GxE native projected feature has zero or invalid variance at block offset {first_var}")
        if first_out is not None:
            raise RuntimeError("GxE native feature output contains a non-finite value")
        host = {name: tab[i].copy() for i, name in enumerate(self._FEATURE_NAMES)}
        host["max_projection_leakage_additive"] = float(np.max(tab[k + 2]))
        host["max_projection_leakage_interaction"] = float(np.max(tab[k + 3]))
        return host

    def _scaled_work(self, W: torch.Tensor, columns: int, sx: np.ndarray, sw: np.ndarray):
        """scale_target_outputs (gxeldcore.cpp:8301-8326): (L, 2C) products -> work_x,
        work_w (L, C) scaled per SNP, F-order views of one pinned host array."""
        W[:, :columns].mul_(torch.as_tensor(sx, device=self._dev)[:, None])
        W[:, columns:].mul_(torch.as_tensor(sw, device=self._dev)[:, None])
        ok = torch.stack([torch.isfinite(W[:, :columns]).all(), torch.isfinite(W[:, columns:]).all()]).cpu()
        if not bool(ok[0]):
            raise RuntimeError("GxE native work_x contains a non-finite value")
        if not bool(ok[1]):
            raise RuntimeError("GxE native work_w contains a non-finite value")
        wt = W.t().contiguous()  # (2C, L): column-major (L, 2C)
        host = torch.empty(wt.shape, dtype=torch.float64, pin_memory=True)
        host.copy_(wt)
        both = host.numpy().T
        return both[:, :columns], both[:, columns:]

    # ------------------------------------------------------------------ API
    def close(self) -> None:
        with self._lock:
            self._close_internal()

    def _close_internal(self) -> None:
        if self._g is not None:
            self._g.close()
            self._g = None
        if self._map is not None:
            self._map.close()
            self._map = None
        for fd in self._fds:
            try:
                os.close(fd)
            except OSError:
                pass
        self._fds = []
        self._closed = True

    def __del__(self):
        try:
            self._close_internal()
        except Exception:
            pass

    def info(self) -> dict:
        with self._lock:
            self._ensure_open()
            return {
                "n_total": self._n_total, "m_total": self._m_total, "n_selected": self._n, "q_rank": self._q,
                "ddof": self._ddof, "decode_threads": self._decode_threads, "blas_threads": self._blas_threads,
                "max_workspace_bytes": self._max_ws, "target_panel_columns": self._tpc,
                "projected_target_full_width": True,
                "strict_feature_moment_verification": self._strict,
                "feature_moment_integrity_mode": ("strict_duplicate" if self._strict
                                                  else "deterministic_disjoint_output_tiled_gemm"),
                "repaired_gemm_output_columns": 0, "retried_gemm_input_mutations": 0,
                "environment_mean": self._env_mean, "environment_variance": self._env_var,
                "max_q_gram_error": self._q_gram_err,
            }

    def feature_block(self, blk_start, blk_end, eps_var=1.0e-10, require_missing_free=True) -> dict:
        """DirectContext::feature_block (gxeldcore.cpp:7349-7416)."""
        with self._lock:
            blk_start, blk_end = int(blk_start), int(blk_end)
            self._validate_block(blk_start, blk_end)
            eps_var = float(eps_var)
            if not (eps_var > 0.0) or not math.isfinite(eps_var):
                raise RuntimeError("GxE native eps_var must be positive and finite")
            self._check_files()
            L = blk_end - blk_start
            copies = 8 if self._strict else 4
            self._ensure_workspace(self._n * L + copies * self._q * L + 11 * L, "feature block")
            missing = self._missing(blk_start, blk_end, bool(require_missing_free))
            res = self._features(blk_start, blk_end, eps_var)
            self._check_files()
            res.update({
                "repaired_feature_moment_columns": 0, "repaired_additive_moment_columns": 0,
                "strict_feature_moment_verification": self._strict,
                "feature_moment_integrity_mode": ("strict_duplicate" if self._strict
                                                  else "deterministic_disjoint_output_tiled_gemm"),
                "missing_genotype_calls": missing,
            })
            return res

    def source_block(self, blk_start, blk_end, scale_x, scale_w, sqrt_annotation, probes,
                     require_missing_free=True):
        """DirectContext::source_block (gxeldcore.cpp:7418-7576): (source_x, source_w,
        missing), two F-order (N, V) views of one (N, 2V) array."""
        with self._lock:
            blk_start, blk_end = int(blk_start), int(blk_end)
            self._validate_block(blk_start, blk_end)
            self._check_files()
            L = blk_end - blk_start
            z = np.asarray(probes, dtype=np.float64)
            sx, sw, a = self._vec(scale_x), self._vec(scale_w), self._vec(sqrt_annotation)
            if z.ndim != 2:
                raise RuntimeError("GxE native source input shape mismatch")
            V = int(z.shape[1])
            if V <= 0 or z.shape[0] != L or sx.size != L or sw.size != L or a.size != L:
                raise RuntimeError("GxE native source input shape mismatch")
            n, p = self._n, self._q
            self._ensure_workspace(n * L + 2 * L * V + 2 * n * V + 2 * p * V + L * V + 3 * L, "source block")
            z = np.array(z, dtype=np.float64)
            missing = self._missing(blk_start, blk_end, bool(require_missing_free))
            bad_scale = ~(np.isfinite(sx) & np.isfinite(sw) & (sx > 0.0) & (sw > 0.0) & np.isfinite(a) & (a >= 0.0))
            bad_probe = ~np.isfinite(z).all(axis=1)
            j_scale = int(np.argmax(bad_scale)) if bad_scale.any() else L
            j_probe = int(np.argmax(bad_probe)) if bad_probe.any() else L
            if j_scale < L and j_scale <= j_probe:
                raise RuntimeError("GxE native source scale/annotation is invalid")
            if j_probe < L:
                raise RuntimeError("GxE native source probe contains a non-finite value")
            za = z * a[:, None]
            coef = np.hstack([za * sx[:, None], za * sw[:, None]]) * self._neg_r[blk_start:blk_end, None]
            # each coefficient column brought to [1/2, 1) by a power of two (exact) and the
            # output scaled back in two exact power-of-two factors (one rounding at most)
            colmax = np.abs(coef).max(axis=0)
            expo = np.where(colmax > 0.0, np.frexp(np.where(colmax > 0.0, colmax, 1.0))[1], 0).astype(np.int64)
            coef = np.ldexp(coef, -expo[None, :])
            out = engine.xc(self._g, torch.as_tensor(coef, device=self._dev), blk_start, blk_end, limbs=LIMBS)
            half = expo // 2
            for part in (half, expo - half):
                out.mul_(torch.as_tensor(np.ldexp(1.0, part), device=self._dev)[:, None])
            out[V:].mul_(self._env_d[None, :])
            self._project(out)
            if not bool(torch.isfinite(out).all().item()):
                raise RuntimeError("GxE native fused source contains a non-finite value")
            # the returned arrays live in a fresh pinned host tensor: one DMA, no staging copy
            host = torch.empty(out.shape, dtype=torch.float64, pin_memory=True)
            host.copy_(out)
            self._check_files()
            fused = host.numpy().T  # (N, 2V), F-order
            return fused[:, :V], fused[:, V:], missing

    def prepare_projected_sources(self, sources, tolerance=1.0e-10) -> ProjectedPanel:
        """DirectContext::prepare_projected_sources (gxeldcore.cpp:7578-7675)."""
        with self._lock:
            self._ensure_open()
            self._check_files()
            s = np.asarray(sources, dtype=np.float64)
            if s.ndim != 2 or s.shape[1] <= 0 or s.shape[0] != self._n:
                raise RuntimeError("GxE native projected source shape mismatch")
            tolerance = float(tolerance)
            if not (tolerance >= 0.0) or not math.isfinite(tolerance):
                raise RuntimeError("GxE native projected source tolerance must be finite and nonnegative")
            C = int(s.shape[1])
            self._ensure_workspace(2 * self._n * C + self._q * C, "projected source preparation")
            if not bool(np.isfinite(s).all()):
                raise RuntimeError("GxE native projected sources contains a non-finite value")
            panel = torch.empty((2 * C, self._n), dtype=torch.float64, device=self._dev)
            panel[:C].copy_(self._upload_panel(s))
            leakage = self._leakage(panel[:C])
            if leakage > tolerance:
                raise RuntimeError("GxE native source is not in the projected fixed-effect complement: "
                                   f"leakage={leakage:f}, tolerance={tolerance:f}")
            # a float64 reprojection seals the panel; e . S serves the W-left products
            self._project(panel[:C])
            torch.mul(panel[:C], self._env_d[None, :], out=panel[C:])
            if not bool(torch.isfinite(panel[C:]).all().item()):
                raise RuntimeError("GxE native environment-weighted projected sources contains a non-finite value")
            self._check_files()
            return ProjectedPanel._seal(self._id, self._n, C, leakage, panel)

    def validate_projected_sources(self, sources, tolerance=1.0e-10) -> float:
        """DirectContext::validate_projected_sources (gxeldcore.cpp:7677-7680)."""
        return self.prepare_projected_sources(sources, tolerance).leakage

    def target_block(self, blk_start, blk_end, scale_x, scale_w, sources, require_missing_free=True):
        """DirectContext::target_block (gxeldcore.cpp:7682-7826): an arbitrary source panel
        S is projected first; (work_x, work_w, missing, largest per-panel leakage)."""
        with self._lock:
            blk_start, blk_end = int(blk_start), int(blk_end)
            self._validate_block(blk_start, blk_end)
            self._check_files()
            L = blk_end - blk_start
            s = np.asarray(sources, dtype=np.float64)
            sx, sw = self._vec(scale_x), self._vec(scale_w)
            if s.ndim != 2 or s.shape[1] <= 0 or s.shape[0] != self._n or sx.size != L or sw.size != L:
                raise RuntimeError("GxE native target input shape mismatch")
            C = int(s.shape[1])
            n, p = self._n, self._q
            width = min(self._tpc, C)
            self._ensure_workspace(n * L + n * C + 2 * L + 2 * L * C + 2 * n * width + p * width, "target block")
            self._validate_scales(sx, sw)
            if not bool(np.isfinite(s).all()):
                raise RuntimeError("GxE native target sources contains a non-finite value")
            missing = self._missing(blk_start, blk_end, bool(require_missing_free))
            panel = torch.empty((2 * C, n), dtype=torch.float64, device=self._dev)
            panel[:C].copy_(self._upload_panel(s))
            leakage = self._leakage(panel[:C], width)
            self._project(panel[:C])
            torch.mul(panel[:C], self._env_d[None, :], out=panel[C:])
            W = self._products(blk_start, blk_end, panel)
            work_x, work_w = self._scaled_work(W, C, sx, sw)
            self._check_files()
            return work_x, work_w, missing, leakage

    def target_projected_block(self, blk_start, blk_end, scale_x, scale_w, sources, require_missing_free=True):
        """DirectContext::target_projected_block / target_projected_panel
        (gxeldcore.cpp:7827-7837, :8341-8470): one traversal over the sealed
        [S, e . S] panel; (work_x, work_w, missing, the panel's leakage)."""
        with self._lock:
            blk_start, blk_end = int(blk_start), int(blk_end)
            self._validate_block(blk_start, blk_end)
            self._check_files()
            self._panel_handle(sources)
            L = blk_end - blk_start
            sx, sw = self._vec(scale_x), self._vec(scale_w)
            if sx.size != L or sw.size != L:
                raise RuntimeError("GxE native target input shape mismatch")
            C = sources._columns
            self._ensure_workspace(self._n * L + 2 * L + 4 * L * C, "projected target block")
            self._validate_scales(sx, sw)
            missing = self._missing(blk_start, blk_end, bool(require_missing_free))
            W = self._products(blk_start, blk_end, sources._data)
            work_x, work_w = self._scaled_work(W, C, sx, sw)
            self._check_files()
            return work_x, work_w, missing, sources._leakage

    def phenotype_score_block(self, blk_start, blk_end, phenotype, eps_var=1.0e-10, require_missing_free=False):
        """DirectContext::phenotype_score_block (gxeldcore.cpp:7839-8029): the block's
        features and the scaled scores X^T P y, W^T P y of a sealed phenotype panel."""
        with self._lock:
            blk_start, blk_end = int(blk_start), int(blk_end)
            self._validate_block(blk_start, blk_end)
            self._panel_handle(phenotype)
            eps_var = float(eps_var)
            if not (eps_var > 0.0) or not math.isfinite(eps_var):
                raise RuntimeError("GxE native phenotype-score eps_var must be positive and finite")
            self._check_files()
            L = blk_end - blk_start
            C = phenotype._columns
            copies = 8 if self._strict else 4
            self._ensure_workspace(self._n * L + copies * self._q * L + 11 * L + 4 * L * C, "phenotype-score block")
            missing = self._missing(blk_start, blk_end, bool(require_missing_free))
            res = self._features(blk_start, blk_end, eps_var)
            W = self._products(blk_start, blk_end, phenotype._data)
            score_x, score_w = self._scaled_work(W, C, res["scale_x"], res["scale_w"])
            self._check_files()
            res.update({
                "score_x": score_x, "score_w": score_w,
                "phenotype_projection_leakage": phenotype._leakage,
                "repaired_feature_moment_columns": 0,
                "missing_genotype_calls": missing,
            })
            return res


# --------------------------------------------------------------------------- probes
def numpy_philox_rademacher_block(keys, rows, threads) -> np.ndarray:
    """gxeldcore.numpy_philox_rademacher_block (gxeldcore.cpp:10070-10136): an F-order
    (rows, probes) FP64 block of +-1 probes, column p the stream NumPy's
    Generator(Philox(key=keys[p])).integers(0, 2, rows, dtype=int8) gives (2 x - 1).
    The streams are generated on the GPU (gpe_philox_rademacher, one counter-based
    thread per Philox block)."""
    threads = int(threads)
    if threads <= 0:
        raise RuntimeError("Protected GxE GEMM threads must be positive")
    if threads > len(os.sched_getaffinity(0)):
        raise RuntimeError("Protected GxE GEMM threads exceed the OpenMP/CPU-affinity limit")
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
    rows = int(rows)
    if k.ndim != 2 or rows <= 0 or k.shape[0] <= 0 or k.shape[1] != 2:
        raise RuntimeError("Native Philox key matrix must be probes-by-two")
    P = int(k.shape[0])
    starts = np.zeros(1, dtype=np.int64)
    z = torch.empty((rows, P), dtype=torch.int8, device=engine.default_device())
    engine.call("gpe_philox_rademacher", int(k.ctypes.data), int(starts.ctypes.data), 1, rows, P, engine._ptr(z), P,
                engine._stream())
    return np.asfortranarray(z.to(torch.float64).cpu().numpy())
