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


def genomewide_ldscore(geno: engine.Genotype, nvecs: int, seed: int, annot: np.ndarray | None = None,
                       step_size: int = 1000, ddof: int = 1, vtiles: list[int] | None = None,
                       rand_dist: str = "rademacher", eps: float = 1e-10, limbs: int = 5, mc: bool = False,
                       stream=None) -> LDScoreResult:
    """The prepared ``geno`` must have been prepared with the same ddof (and the
    impute mode / seed of the run)."""
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
    v_start = 0
    guards = [0, 0]
    for Vt in vtiles:
        xz = torch.zeros((K * Vt, N), dtype=torch.float64, device=dev)
        engine.phase1_xz(geno, 0, M, step_size, annot_d, inv, v_start, Vt, seed, xz, rand_dist=rand_dist,
                         limbs=limbs, stream=stream)
        gc = engine.phase2_accum(geno, 0, M, inv, Vt, xz, meansq, n_denom, mc_aggregate_sumsq=mc_agg,
                                 mc_per_snp_sumsq=mc_ps, limbs=limbs, stream=stream, want_guards=True)
        guards[0] += gc[0]
        guards[1] += gc[1]
        v_start += Vt
    torch.cuda.synchronize()
    meansq_h = meansq.cpu().numpy()
    nsnps_bin = annot.sum(axis=0, dtype=np.float64)
    ld = meansq_h / float(nvecs) - (nsnps_bin / float(N - ddof))[None, :]
    return LDScoreResult(ld, meansq_h, inv.cpu().numpy(), int(nvecs), nsnps_bin,
                         None if mc_ps is None else mc_ps.cpu().numpy(), tuple(guards))


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
