"""Drop-in mirror of the reference's ``summit.winldcore`` (exact windowed LD scores).

``compute_windowed_ld_chr`` keeps the keyword contract of
src/native/winldcore.cpp:964-994 and returns an (m, B) Fortran FP64 array.
The banded Gram runs on int8 MFMA with exact integer cross products
(``gpe_windowed_ld``); panel caching (panel_cols, cache_mb) has no role on a
device-resident matrix and those arguments are accepted and ignored.
With covariates (``C`` (N, p) and ``R`` (p, N) both given) every SNP column is
projected and restandardized as the reference's panels are
(src/native/winldcore.cpp:341-372), with corr_dim = N - p
(``gpe_windowed_ld_cov``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import engine
from .gwldcore import _geno, _to_dev, get_max_threads, set_num_threads, set_verbose  # noqa: F401

_PROGRESS = {"total": 0, "done": 0, "active": False}


def get_progress_total() -> int:
    return _PROGRESS["total"]


def get_progress_done() -> int:
    return _PROGRESS["done"]


def progress_active() -> bool:
    return _PROGRESS["active"]


def windowed_ld(g: engine.Genotype, snp_begin: int, snp_end: int, bp=None, ld_wind_kb: float | None = None,
                annot=None, window_snps: int = 0, C=None, R=None, side: int = 0) -> np.ndarray:
    """Device-level call: (m, B) FP64 windowed LD scores of SNPs [snp_begin, snp_end);
    with C (N, p) and R (p, N) the panels are projected off the covariates; side 1
    (2) sums only the window SNPs before (after) each SNP."""
    m = snp_end - snp_begin
    if annot is None:
        annot = np.ones((m, 1))
    annot = np.ascontiguousarray(np.asarray(annot, dtype=np.float64))
    if annot.ndim != 2 or annot.shape[0] != m:
        raise RuntimeError("annot_chr shape mismatch in compute_windowed_ld_chr")
    B = annot.shape[1]
    bp_arr = None
    if window_snps <= 0:
        bp_arr = np.ascontiguousarray(np.asarray(bp, dtype=np.int64))
        if bp_arr.shape[0] != m:
            raise RuntimeError("bp length mismatch in compute_windowed_ld_chr")
    a_d = torch.from_numpy(annot).to(engine.default_device())
    out = torch.empty((m, B), dtype=torch.float64, device=a_d.device)
    window_bp = float(ld_wind_kb) * 1000.0 if ld_wind_kb is not None else 0.0
    bp_ptr = 0 if bp_arr is None else int(bp_arr.ctypes.data)
    if C is not None and R is not None:
        Cn = np.asarray(C.cpu().numpy() if isinstance(C, torch.Tensor) else C, dtype=np.float64)
        Rn = np.asarray(R.cpu().numpy() if isinstance(R, torch.Tensor) else R, dtype=np.float64)
        if Cn.ndim != 2 or Rn.ndim != 2 or Cn.shape[0] != g.n or Rn.shape != (Cn.shape[1], g.n):
            raise RuntimeError("C/R shape mismatch in compute_windowed_ld_chr")
        p = Cn.shape[1]
        c_d = torch.from_numpy(np.ascontiguousarray(Cn.T)).to(a_d.device)  # column k at k * N
        r_d = torch.from_numpy(np.ascontiguousarray(Rn)).to(a_d.device)    # row k at k * N
        engine.call("gpe_windowed_ld_cov", g.handle, int(snp_begin), int(snp_end), bp_ptr, window_bp,
                    int(window_snps), engine._ptr(a_d), B, engine._ptr(c_d), g.n, engine._ptr(r_d), g.n, p,
                    engine._ptr(out), engine._stream())
        return out.cpu().numpy()
    if side:
        engine.call("gpe_windowed_ld_side", g.handle, int(snp_begin), int(snp_end), bp_ptr, window_bp,
                    int(window_snps), int(side), engine._ptr(a_d), B, engine._ptr(out), engine._stream())
        return out.cpu().numpy()
    engine.call("gpe_windowed_ld", g.handle, int(snp_begin), int(snp_end), bp_ptr, window_bp, int(window_snps),
                engine._ptr(a_d), B, engine._ptr(out), engine._stream())
    return out.cpu().numpy()


def windowed_ld_multi(g: engine.Genotype, chroms, ld_wind_kb: float | None = None, annot=None,
                      window_snps: int = 0) -> np.ndarray:
    """Every chromosome in one banded-Gram launch.  chroms: [(snp_begin,
    snp_end, bp or None), ...] sorted and non-overlapping; returns (rows, B)
    FP64 for SNPs [chroms[0][0], chroms[-1][1]) (rows between chromosomes 0),
    equal to one ``windowed_ld`` call per chromosome up to FP64 round-off (tile
    pairs add into each SNP's score by FP64 atomics, in no fixed order)."""
    bounds = np.ascontiguousarray(np.asarray([[int(c[0]), int(c[1])] for c in chroms], dtype=np.int64).reshape(-1))
    if bounds.size == 0:
        raise RuntimeError("at least one chromosome is required")
    row0, rows = int(bounds[0]), int(bounds[-1] - bounds[0])
    if annot is None:
        annot = np.ones((rows, 1))
    annot = np.ascontiguousarray(np.asarray(annot, dtype=np.float64))
    if annot.ndim != 2 or annot.shape[0] != rows:
        raise RuntimeError("annot_chr shape mismatch in compute_windowed_ld_chr")
    bp_all = None
    if window_snps <= 0:
        bp_all = np.zeros(rows, dtype=np.int64)
        for s0, s1, bp in chroms:
            bp = np.asarray(bp, dtype=np.int64)
            if bp.shape[0] != s1 - s0:
                raise RuntimeError("bp length mismatch in compute_windowed_ld_chr")
            bp_all[s0 - row0:s1 - row0] = bp
    a_d = torch.from_numpy(annot).to(engine.default_device())
    out = torch.empty((rows, annot.shape[1]), dtype=torch.float64, device=a_d.device)
    window_bp = float(ld_wind_kb) * 1000.0 if ld_wind_kb is not None else 0.0
    engine.call("gpe_windowed_ld_multi", g.handle, int(bounds.size // 2), int(bounds.ctypes.data),
                0 if bp_all is None else int(bp_all.ctypes.data), window_bp, int(window_snps), engine._ptr(a_d),
                annot.shape[1], engine._ptr(out), engine._stream())
    return out.cpu().numpy()


def compute_windowed_ld_chr(bed_prefix, fam_path, chr_start, chr_end, bp, annot_chr, ld_wind_kb, chunk_size,
                            row_sel=None, C=None, R=None, impute_mode="mean", impute_seed=None, panel_cols=0,
                            cache_mb=-1):
    """src/native/winldcore.cpp:783-931."""
    chr_start, chr_end = int(chr_start), int(chr_end)
    if chr_end < chr_start:
        raise RuntimeError("chr_end must be >= chr_start")
    m = chr_end - chr_start
    if m <= 0:
        raise RuntimeError("Chromosome slice is empty")
    if int(chunk_size) <= 0:
        raise RuntimeError("chunk_size must be > 0")
    wbp = float(ld_wind_kb) * 1000.0
    if not np.isfinite(wbp) or wbp <= 0.0:
        raise RuntimeError("ld_wind_kb must define a finite positive BP window")
    g, _, m_total = _geno(bed_prefix, fam_path, row_sel, 0, impute_mode, impute_seed)
    if chr_end > m_total:
        raise RuntimeError("chr_end exceeds #SNPs in BIM")
    projecting = C is not None and R is not None
    if projecting and g.n - np.asarray(C).shape[1] <= 1:
        raise RuntimeError("Too few residual dimensions for unbiased windowed r^2 correction")
    _PROGRESS.update(total=1, done=0, active=True)
    try:
        out = windowed_ld(g, chr_start, chr_end, bp, ld_wind_kb, annot_chr, C=C if projecting else None,
                          R=R if projecting else None)
    finally:
        _PROGRESS.update(done=1, active=False)
    return np.asfortranarray(out)
