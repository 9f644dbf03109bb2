"""Genetic covariance and correlation between two traits from the engine's LD
scores: the cross-trait moment (normal-equation) solve with a delete-one
jackknife, in FP64 on the host.

Restated from SUMMIT's cross-trait estimator:

* the response per SNP, ``y_j = z1*_j z2*_j``, the product of the two traits'
  score statistics on their own effective sample sizes
  (src/summit/sumstats/moments.py:185-236);
* the SNP-specific overlap multiplier ``sqrt(s1_j s2_j)``, ``s_j = N*/(n_j - q - 1)``
  the null variance of ``z*_j``, summed per unit into the overlap mass
  ``C_k = sum a_jk sqrt(s1_j s2_j)`` (src/summit/sumstats/moments.py:108-126,
  src/summit/inference/rgcore.py:463-472, :588-592); with complete per-SNP
  sample sizes the multiplier is one and ``C_k = A_k``;
* per jackknife unit the sums ``A_k``, ``(A^T L)`` and ``A^T y`` and their
  delete-one replicates (rgcore.py:528-545);
* the cross traces ``T_kl = (A^T L)_kl N1* N2* / (A_k A_l)``, zero where the
  masses vanish (src/summit/utils.py:558-571), with the replicate's own mass
  times the full mass (rgcore.py:555-564), symmetrized with the same
  jackknife-weighted pairing as the univariate traces (rgcore.py:566);
* per replicate ``T gamma = sqrt(N1* N2*) (A^T y - c C) / A`` (rgcore.py:1285-1299)
  for a supplied overlap covariance ``c`` on the response's scale (the mean of
  ``y`` at a SNP that tags nothing: zero for disjoint samples); solved with LU,
  least squares when singular (rgcore.py:340-361);
* per-bin ``rg_k = gamma_k / sqrt(v1_k v2_k)`` with ``v`` the two univariate fits'
  per-bin variance components of the same replicate, the total
  ``rg = sum_k gamma_k / sqrt(h2_1 h2_2)`` (rgcore.py:81-91, :1310-1350), and the
  delta-method alternative for the total's SE (rgcore.py:633-658).

Standard errors are the unequal-unit delete-one pseudovalue SEs of the
univariate fit, on the units of its jackknife design.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .h2 import (JackknifeDesign, H2Fit, _replicates, _solve_each, effective_n_scale, jackknife_se, score_noise_scale,
                 score_z, symmetrize_traces, unit_sums)


def rg_response(beta1, se1, n_obs1, nsamp1: float, beta2, se2, n_obs2, nsamp2: float,
                cov_rank1: int = 0, cov_rank2: int = 0):
    """(y, overlap multiplier, N1*, N2*): the per-SNP cross-trait response
    ``z1* z2*`` (NaN where either statistic is undefined), ``sqrt(s1 s2)`` with
    ``s`` each statistic's null variance, and the two effective sample sizes."""
    z1 = score_z(beta1, se1, n_obs1, nsamp1, cov_rank1)
    z2 = score_z(beta2, se2, n_obs2, nsamp2, cov_rank2)
    if z1.shape != z2.shape:
        raise ValueError(f"trait shapes differ: {z1.shape} vs {z2.shape}")
    y = z1 * z2
    y[~np.isfinite(y)] = np.nan
    s1 = score_noise_scale(n_obs1, nsamp1, cov_rank1, shape=z1.shape)
    s2 = score_noise_scale(n_obs2, nsamp2, cov_rank2, shape=z2.shape)
    with np.errstate(invalid="ignore"):
        overlap = np.sqrt(s1 * s2)
    return y, overlap, effective_n_scale(nsamp1, cov_rank1), effective_n_scale(nsamp2, cov_rank2)


def component_rg(gamma, v1, v2) -> np.ndarray:
    """``gamma / sqrt(v1 v2)`` where both variances are finite and positive, else NaN."""
    g, a, b = np.broadcast_arrays(np.asarray(gamma, dtype=np.float64), np.asarray(v1, dtype=np.float64),
                                  np.asarray(v2, dtype=np.float64))
    out = np.full(g.shape, np.nan)
    ok = np.isfinite(g) & np.isfinite(a) & np.isfinite(b) & (a > 0) & (b > 0)
    out[ok] = g[ok] / np.sqrt(a[ok] * b[ok])
    return out


@dataclass
class RGFit:
    gamma: np.ndarray         # (K,) per-bin genetic covariance components
    gamma_se: np.ndarray
    rg: np.ndarray            # (K,) per-bin genetic correlations
    rg_se: np.ndarray
    gamma_total: float        # total genetic covariance
    gamma_total_se: float
    rg_total: float           # total genetic correlation
    rg_total_se: float
    c: float                  # the overlap covariance used
    se_method: str
    n1_scale: float
    n2_scale: float
    nsnps: int
    gamma_reps: np.ndarray    # (U + 1, K) replicate solutions, full estimate last
    rg_reps: np.ndarray       # (U + 1, K)
    rg_total_reps: np.ndarray  # (U + 1,)


def _delta_rg_total_se(gamma_tot, h2_1, h2_2, sizes) -> float:
    """SE of the total rg from the jackknife of its first-order expansion around
    the full estimate in (gamma, h2_1, h2_2)."""
    g, v1, v2 = float(gamma_tot[-1]), float(h2_1[-1]), float(h2_2[-1])
    if not (np.isfinite(g) and np.isfinite(v1) and np.isfinite(v2) and v1 > 0 and v2 > 0):
        return np.nan
    r = g / np.sqrt(v1 * v2)
    lin = r + (gamma_tot - g) / np.sqrt(v1 * v2) - 0.5 * r * (h2_1 - v1) / v1 - 0.5 * r * (h2_2 - v2) / v2
    lin[-1] = r
    return float(jackknife_se(lin[:, None], sizes)[0])


def fit_rg(ldscores, annot, y, n1_scale: float, n2_scale: float, design: JackknifeDesign,
           h2_fit1: H2Fit, h2_fit2: H2Fit, c: float = 0.0, overlap=None, se_method: str = "jackknife") -> RGFit:
    """Solve the cross-trait moment equations on SNPs with a finite response.

    ldscores, annot: (M, K) as for the univariate fit; y: (M,) ``rg_response``'s
    response; n1_scale, n2_scale: N1*, N2*; h2_fit1, h2_fit2: the two traits'
    univariate fits on the same SNPs and jackknife design (their replicate
    variance components scale each replicate's rg); c: the overlap covariance
    (zero when the two GWAS share no samples); overlap: the per-SNP overlap
    multiplier (``rg_response``; ``None`` = one); se_method: ``"jackknife"`` or
    ``"delta"`` (the total's SE from the expansion in gamma and both h2)."""
    L = np.asarray(ldscores, dtype=np.float64)
    A = np.asarray(annot, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    if L.ndim == 1:
        L = L[:, None]
    if A.ndim == 1:
        A = A[:, None]
    M, K = A.shape
    if L.shape != (M, K) or y.shape != (M,):
        raise ValueError(f"A/L shape mismatch: {A.shape} vs {L.shape}")
    if design.nsnps != M:
        raise ValueError(f"JackknifeDesign was not built on this trace axis. Expected {M}, got {design.nsnps}.")
    se_method = str(se_method).strip().lower()
    if se_method not in ("jackknife", "delta"):
        raise ValueError("se_method must be one of {'jackknife','delta'}")
    n1, n2 = float(n1_scale), float(n2_scale)
    if not (np.isfinite(n1) and n1 > 0.0 and np.isfinite(n2) and n2 > 0.0):
        raise RuntimeError(f"Invalid rg n_scale pair: n1_scale={n1_scale}, n2_scale={n2_scale}")
    c = float(c)
    if not np.isfinite(c):
        raise ValueError("the overlap covariance c must be finite.")
    active = np.isfinite(y)
    if not active.any():
        raise ValueError("No SNPs remain after filtering non-finite rg summary moments.")
    U = design.nunit
    for f in (h2_fit1, h2_fit2):
        if f.sigma_reps is None or f.sigma_reps.shape[0] != U + 1 or f.sigma_reps.shape[1] < K:
            raise ValueError("the univariate fits must come from fit_h2 on the same jackknife design and bins.")
    m_u, Ak_u, _, AL_u, Ay_u = unit_sums(A, L, y, active, design)
    Ak_rep = _replicates(Ak_u)
    AL_rep = _replicates(AL_u)
    Ay_rep = _replicates(Ay_u)
    if overlap is None:
        Cm_rep = Ak_rep
    else:
        ov = np.broadcast_to(np.asarray(overlap, dtype=np.float64), (M,))
        Cm_u = np.zeros((U, K))
        for u in range(U):
            s, e = int(design.starts[u]), int(design.ends[u])
            act = active[s:e]
            Cm_u[u] = A[s:e][act].T @ ov[s:e][act]
        Cm_rep = _replicates(Cm_u)
    denom = Ak_rep[:, :, None] * Ak_rep[-1][None, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        lhs = AL_rep * (n1 * n2) / denom
    lhs = np.where((denom > 0) & np.isfinite(denom), lhs, 0.0)
    lhs = symmetrize_traces(lhs, m_u)
    sqn = float(np.sqrt(n1 * n2))
    with np.errstate(divide="ignore", invalid="ignore"):
        rhs = (Ay_rep - c * Cm_rep) * sqn / Ak_rep
    rhs[~np.isfinite(rhs) | ~np.isfinite(Ak_rep) | (Ak_rep <= 0.0)] = np.nan
    gamma_reps = _solve_each(lhs, rhs)
    rg_reps = component_rg(gamma_reps, h2_fit1.sigma_reps[:, :K], h2_fit2.sigma_reps[:, :K])
    gamma_tot = np.nansum(gamma_reps, axis=1)
    rg_tot = component_rg(gamma_tot, h2_fit1.reps_h2, h2_fit2.reps_h2)
    rg_se = jackknife_se(rg_reps, m_u)
    rg_total_se = float(jackknife_se(rg_tot[:, None], m_u)[0])
    rg_full = rg_reps[-1].copy()
    if se_method == "delta":
        rg_total_se = _delta_rg_total_se(gamma_tot, h2_fit1.reps_h2, h2_fit2.reps_h2, m_u)
        if K == 1 and np.isfinite(rg_tot[-1]):
            rg_full[0] = rg_tot[-1]
            rg_se = np.array([rg_total_se])
    return RGFit(gamma=gamma_reps[-1], gamma_se=jackknife_se(gamma_reps, m_u), rg=rg_full, rg_se=rg_se,
                 gamma_total=float(gamma_tot[-1]), gamma_total_se=float(jackknife_se(gamma_tot[:, None], m_u)[0]),
                 rg_total=float(rg_tot[-1]), rg_total_se=rg_total_se, c=c, se_method=se_method,
                 n1_scale=n1, n2_scale=n2, nsnps=int(active.sum()), gamma_reps=gamma_reps, rg_reps=rg_reps,
                 rg_total_reps=rg_tot)
