"""Synthetic GWAS on a device-resident genotype matrix (a workload generator,
no reference counterpart): a phenotype with known, annotation-partitioned
heritability and its per-SNP least-squares effects, formed with the engine's
own products (``gpe_xc_f64`` for X beta, ``gpe_xtv_f64`` for X^T y).  It feeds
the heritability solve (``summit_amd.h2``) end to end."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import engine


@dataclass
class Gwas:
    beta: np.ndarray       # (M,) least-squares effect of each standardized SNP
    se: np.ndarray         # (M,)
    n: int                 # samples
    h2_bins: np.ndarray    # (K,) heritability put on each annotation bin
    h2_total: float        # realized var(X beta) / var(y)


def gwas(g: engine.Genotype, annot: np.ndarray, h2_bins, seed: int, limbs: int = 6) -> Gwas:
    """Phenotype y = X beta + e on the prepared handle's standardized genotypes
    (ddof 1, mean imputation: ||x_j||^2 = nobs_j - 1), with beta_j ~ N(0, sum_k
    a_jk h2_k / A_k) (A_k = sum_j a_jk) rescaled so var(X beta) / var(y) is the
    total; then b_j = x_j^T y / ||x_j||^2 and its OLS standard error."""
    if g.prepared_with is None or g.prepared_with[0] != 1:
        raise RuntimeError("gwas needs a handle prepared with ddof=1")
    rng = np.random.default_rng(seed)
    A = np.asarray(annot, dtype=np.float64)
    h2b = np.asarray(h2_bins, dtype=np.float64)
    M, N = g.m, g.n
    if A.shape != (M, h2b.size):
        raise RuntimeError("annotation must be (M, K) with one heritability per bin")
    per_snp = A @ (h2b / A.sum(axis=0))
    dev = engine.default_device()
    beta = torch.from_numpy((rng.standard_normal(M) * np.sqrt(per_snp))[:, None]).to(dev)
    gv = engine.xc(g, beta, limbs=limbs)[0].cpu().numpy()
    h2 = float(h2b.sum())
    gv = (gv - gv.mean()) * np.sqrt(h2 / gv.var())
    e = rng.standard_normal(N)
    e = (e - e.mean()) * np.sqrt((1.0 - h2) / e.var())
    y = gv + e
    y -= y.mean()
    xty = engine.xtv(g, torch.from_numpy(y[None, :]).to(dev), limbs=limbs)[:, 0].cpu().numpy()
    st = g.stats()
    nobs = st["nobs"].astype(np.int64)
    # a column with two or more distinct observed values standardizes to norm^2 nobs - 1;
    # any other column is identically zero
    poly = (nobs > 1) & (nobs * st["sumsq"] != st["sum"] * st["sum"])
    xx = np.where(poly, nobs - 1.0, 0.0)
    with np.errstate(divide="ignore", invalid="ignore"):
        b = np.where(xx > 0, xty / xx, np.nan)
        resid = float(y @ y) - b * b * xx
        se = np.sqrt(resid / (N - 2) / xx)
    return Gwas(beta=b, se=se, n=N, h2_bins=h2b, h2_total=float(gv.var() / y.var()))
