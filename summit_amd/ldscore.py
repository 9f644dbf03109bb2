"""Genome-wide / partitioned stochastic LD scores on one GPU.

Mirrors ``GenomewideLDScore._compute_ldscore`` (src/summit/ldscore/gw_ldscore.py:1753-2223)
for BED input without covariates: residual scales, then per probe tile one
phase-1 traversal (X.Z over every SNP block, probe streams seeded per block)
and one phase-2 traversal (X^T(X.Z) with the squared-norm reduction), then the
null subtraction ``meansq / nvecs - M_k / (N - ddof)``.  The genotype matrix
stays resident in HBM across both traversals and every probe tile.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import engine


def impute_seed_from_root(root_seed: int) -> int:
    """src/summit/ldscore/gw_ldscore.py:1359"""
    return (int(root_seed) ^ 0xA24BAED4963EE407) & 0xFFFFFFFFFFFFFFFF


def bounded_vtile_sizes(nvecs: int, vmax: int, granularity: int = 64) -> list[int]:
    """Probe tiles under a memory ceiling (src/summit/ldscore/gw_ldscore.py:687-697)."""
    nvecs, vmax, granularity = int(nvecs), int(vmax), int(granularity)
    if nvecs <= 0 or vmax <= 0 or granularity <= 0:
        raise ValueError("nvecs, vmax, and granularity must be positive integers.")
    width = min(nvecs, vmax)
    if width >= granularity:
        width = max(granularity, (width // granularity) * granularity)
    return [min(width, nvecs - start) for start in range(0, nvecs, width)]


@dataclass
class LDScoreResult:
    ldscore: np.ndarray        # (M, K) FP64, null-subtracted
    meansq_accum: np.ndarray   # (M, K) FP64 raw accumulators (sum over probes)
    inv: np.ndarray            # (M,) residual scales
    nvecs: int
    nsnps_bin: np.ndarray
    mc_per_snp_sumsq: np.ndarray | None = None
    guard_counts: tuple = (0, 0)
    mc_var: np.ndarray | None = None  # per-SNP MC variance of ldscore (mc=True)


def traversal_plan(vtiles: list[int], width: int | None) -> list[tuple[int, int, int]]:
    """(v_start, v_count, probe_tile) per traversal: consecutive V-tiles of one
    size merge into traversals of up to ``width`` probe columns (None: one
    traversal per V-tile)."""
    plan, v0, i = [], 0, 0
    while i < len(vtiles):
        P, k = int(vtiles[i]), 1
        if width:
            while i + k < len(vtiles) and int(vtiles[i + k]) == P and (k + 1) * P <= width:
                k += 1
        plan.append((v0, k * P, P))
        v0 += k * P
        i += k
    return plan


def genomewide_ldscore(geno: engine.Genotype, nvecs: int, seed: int, annot: np.ndarray | None = None,
                       step_size: int = 1000, ddof: int = 1, vtiles: list[int] | None = None,
                       rand_dist: str = "rademacher", eps: float = 1e-10, limbs: int = 5, mc: bool = False,
                       stream=None, traversal_width: int | None = None) -> LDScoreResult:
    """The prepared ``geno`` must have been prepared with the same ddof (and the
    impute mode / seed of the run).  ``vtiles`` are the reference's V-tiles
    (they fix the probe streams); ``traversal_width`` lets one pair of
    traversals serve several equal V-tiles (same estimate, fewer, wider
    traversals, and one probe stream per block and tile in parallel)."""
    if geno.prepared_with is None:
        raise RuntimeError("genotype handle is not prepared")
    M, N = geno.m, geno.n
    dev = engine.default_device()
    if annot is None:
        annot = np.ones((M, 1), dtype=np.float64)
    annot = np.asarray(annot, dtype=np.float64)
    if annot.shape[0] != M:
        raise RuntimeError("annotation rows must equal the SNP count")
    K = annot.shape[1]
    vtiles = [int(nvecs)] if vtiles is None else [int(v) for v in vtiles]
    if sum(vtiles) != nvecs:
        raise RuntimeError("V-tiles must sum to nvecs")
    annot_d = torch.from_numpy(np.ascontiguousarray(annot)).to(dev)
    inv = engine.resvar_inv(geno, 0, M, eps, stream=stream)
    meansq = torch.zeros((M, K), dtype=torch.float64, device=dev)
    mc_agg = torch.zeros(K, dtype=torch.float64, device=dev) if mc else None
    mc_ps = torch.zeros((M, K), dtype=torch.float64, device=dev) if mc else None
    n_denom = N - ddof + 1
    guards = [0, 0]
    for v_start, Vt, tile in traversal_plan(vtiles, traversal_width):
        xz = torch.zeros((K * Vt, N), dtype=torch.float64, device=dev)
        engine.phase1_xz(geno, 0, M, step_size, annot_d, inv, v_start, Vt, seed, xz, rand_dist=rand_dist,
                         limbs=limbs, stream=stream, probe_tile=tile)
        gc = engine.phase2_accum(geno, 0, M, inv, Vt, xz, meansq, n_denom, mc_aggregate_sumsq=mc_agg,
                                 mc_per_snp_sumsq=mc_ps, limbs=limbs, stream=stream, want_guards=True)
        guards[0] += gc[0]
        guards[1] += gc[1]
    torch.cuda.synchronize()
    meansq_h = meansq.cpu().numpy()
    nsnps_bin = annot.sum(axis=0, dtype=np.float64)
    ld = meansq_h / float(nvecs) - (nsnps_bin / float(N - ddof))[None, :]
    mc_ps_h = None if mc_ps is None else mc_ps.cpu().numpy()
    mc_var = None if mc_ps_h is None else mc_variance_from_probe_sums(meansq_h, mc_ps_h, int(nvecs))
    return LDScoreResult(ld, meansq_h, inv.cpu().numpy(), int(nvecs), nsnps_bin, mc_ps_h, tuple(guards), mc_var)


class GraphedTiles:
    """Probe tiles of the genome-wide estimator replayed from one HIP graph.

    One tile = clear X.Z, phase 1 over every SNP block (Rademacher streams of
    V-tile ``v_start`` read on the device), phase 2 accumulating into
    ``meansq``, and ``v_start += tile``: the launches of a tile are captured
    once (after one eager tile sizes every workspace) and each ``replay()``
    runs the next tile.  Results equal eager tiles bit for bit."""

    def __init__(self, geno: engine.Genotype, tile: int, seed: int, annot: torch.Tensor, inv: torch.Tensor,
                 meansq: torch.Tensor, step_size: int = 1000, limbs: int = 5, probe_tile: int | None = None,
                 v_start: int = 0):
        dev = engine.default_device()
        self.g, self.tile, self.seed = geno, int(tile), int(seed)
        self.annot, self.inv, self.meansq = annot, inv, meansq
        self.step_size, self.limbs = int(step_size), int(limbs)
        self.probe_tile = probe_tile
        K = annot.shape[1]
        self.xz = torch.zeros((K * self.tile, geno.n), dtype=torch.float64, device=dev)
        self.v_start = torch.tensor([int(v_start)], dtype=torch.int32, device=dev)
        self.n_denom = geno.n - 1 + 1  # ddof 1
        self.graph = None

    def _tile(self, stream=None) -> None:
        self.xz.zero_()
        engine.phase1_xz(self.g, 0, self.g.m, self.step_size, self.annot, self.inv, 0, self.tile, self.seed,
                         self.xz, limbs=self.limbs, probe_tile=self.probe_tile, v_start_dev=self.v_start,
                         stream=stream)
        engine.phase2_accum(self.g, 0, self.g.m, self.inv, self.tile, self.xz, self.meansq, self.n_denom,
                            limbs=self.limbs, stream=stream)
        engine.add_i32(self.v_start, self.tile, stream=stream)

    def eager(self) -> None:
        """One tile without the graph (also sizes the workspaces)."""
        self._tile()

    def capture(self) -> None:
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, stream=side):
            self._tile(stream=side)
        torch.cuda.current_stream().wait_stream(side)

    def replay(self) -> None:
        if self.graph is None:
            raise RuntimeError("capture() first (after one eager tile)")
        self.graph.replay()


def write_gw_ldscore_table(outpath, snpdf, scores_df, *, snp_var=None, ld_rel=None) -> str:
    """``<outpath>.gw.ldscore.gz``: the SNP columns (CHR SNP BP), V when given, one
    column per annotation and LD_REL_<column> when given; tab-separated, six
    decimals, gzip level 6 (src/summit/ldscore/gw_ldscore.py:1186-1200).
    ``snpdf`` and ``scores_df`` share a row index."""
    import pandas as pd

    parts = [snpdf]
    if snp_var is not None:
        parts.append(pd.DataFrame({"V": snp_var}))
    parts.append(scores_df)
    if ld_rel is not None:
        parts.append(pd.DataFrame(ld_rel, columns=[f"LD_REL_{c}" for c in scores_df.columns]))
    table = pd.concat(parts, axis=1)
    path = f"{outpath}.gw.ldscore.gz"
    table.to_csv(path, index=False, compression={"method": "gzip", "compresslevel": 6}, sep="\t",
                 float_format="%.6f")
    return path


def write_gw_annotation_masses(outpath, masses) -> str:
    """``<outpath>.gw.M``: annotation masses on one tab-separated line, %.10g
    (src/summit/ldscore/gw_ldscore.py:1203-1208)."""
    path = f"{outpath}.gw.M"
    with open(path, "w") as f:
        f.write("\t".join(f"{float(x):.10g}" for x in masses) + "\n")
    return path


def write_gw_delta(outpath, delta, l2cols) -> str:
    """``<outpath>.gw.delta``: the block-level fourth-moment matrix
    delta = mu22 - (1 + 2 rho2) (``compute_block_corrections_binary``), rows and
    columns named by the annotation columns, tab-separated, %.8e
    (src/summit/ldscore/gw_ldscore.py:2343-2349)."""
    import pandas as pd

    path = f"{outpath}.gw.delta"
    pd.DataFrame(np.asarray(delta, dtype=np.float64), index=list(l2cols), columns=list(l2cols)).to_csv(
        path, sep="\t", float_format="%.8e")
    return path


def mc_variance_from_probe_sums(probe_sum, probe_sumsq, nvecs: int, return_info: bool = False):
    """Monte Carlo variance of a probe mean, sample_var(Y_v) / V, from S1 = sum_v Y_v
    and S2 = sum_v Y_v^2 per element: (S2 - S1^2 / V) / (V (V - 1))
    (src/summit/ldscore/gw_ldscore.py:874-944).  A negative centred sum within
    512 eps max(|S2|, |S1^2/V|, 1) is round-off and becomes 0; beyond that, or
    non-finite, the element is NaN.  V < 2 gives NaN everywhere."""
    first = np.asarray(probe_sum)
    second = np.asarray(probe_sumsq)
    if first.shape != second.shape:
        raise ValueError("probe_sum and probe_sumsq must have identical shapes")
    if first.ndim != 2:
        raise ValueError("probe_sum and probe_sumsq must be two-dimensional")
    v = int(nvecs)
    if v < 2:
        res = np.full(first.shape, np.nan)
        info = {"n_roundoff_clipped": np.zeros(first.shape[1], dtype=np.int64),
                "n_numerical_failures": np.full(first.shape[1], first.shape[0], dtype=np.int64)}
        return (res, info) if return_info else res
    eps = np.finfo(first.dtype).eps if np.issubdtype(first.dtype, np.floating) else np.finfo(np.float64).eps
    f = first.astype(np.float64)
    s = second.astype(np.float64)
    corr = f * f / float(v)
    centred = s - corr
    tol = 512.0 * eps * np.maximum(np.maximum(np.abs(s), np.abs(corr)), 1.0)
    finite = np.isfinite(centred) & np.isfinite(tol)
    tiny = finite & (centred < 0.0) & (centred >= -tol)
    bad = ~finite | (centred < -tol)
    centred = np.where(tiny, 0.0, centred)
    centred = np.where(bad, np.nan, centred)
    res = centred / float(v * (v - 1))
    info = {"n_roundoff_clipped": tiny.sum(axis=0).astype(np.int64),
            "n_numerical_failures": bad.sum(axis=0).astype(np.int64)}
    return (res, info) if return_info else res


def integrated_mc_variance_from_probe_sums(per_snp_probe_sum, aggregate_probe_sumsq, nvecs: int,
                                           return_info: bool = False):
    """Sum over SNPs of the per-SNP MC variances from the per-SNP first sums and the
    per-bin aggregate second sums (src/summit/ldscore/gw_ldscore.py:947-995)."""
    first = np.asarray(per_snp_probe_sum)
    second = np.asarray(aggregate_probe_sumsq, dtype=np.float64)
    if first.ndim != 2 or second.shape != (first.shape[1],):
        raise ValueError("aggregate MC sufficient-statistic shapes are inconsistent")
    v = int(nvecs)
    if v < 2:
        res = np.full(second.shape, np.nan)
        info = {"n_roundoff_clipped": np.zeros(second.shape, dtype=np.int64),
                "n_numerical_failures": np.ones(second.shape, dtype=np.int64)}
        return (res, info) if return_info else res
    f = first.astype(np.float64)
    fsq = np.einsum("ij,ij->j", f, f)
    centred = second - fsq / float(v)
    eps = np.finfo(first.dtype).eps if np.issubdtype(first.dtype, np.floating) else np.finfo(np.float64).eps
    tol = 512.0 * eps * np.maximum(np.maximum(np.abs(second), np.abs(fsq / float(v))), 1.0)
    finite = np.isfinite(centred) & np.isfinite(tol)
    tiny = finite & (centred < 0.0) & (centred >= -tol)
    bad = ~finite | (centred < -tol)
    centred = np.where(tiny, 0.0, centred)
    centred = np.where(bad, np.nan, centred)
    res = centred / float(v * (v - 1))
    info = {"n_roundoff_clipped": tiny.astype(np.int64), "n_numerical_failures": bad.astype(np.int64)}
    return (res, info) if return_info else res
