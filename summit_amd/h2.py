"""Heritability from the engine's LD scores: the moment (normal-equation) solve
with a delete-one jackknife, in FP64 on the host.

This is the consumer of the LD scores the engine writes (``.gw.ldscore.gz`` /
``.gw.M``), restated from SUMMIT's univariate estimator:

* the GWAS response per SNP, ``y_j = z*_j^2 - (N*/(n_j - q - 1) - 1)`` with the
  score statistic ``z*_j = sqrt(N*) b_j / sqrt(b_j^2 + (n_j - q - 2) se_j^2)``
  and ``N* = N - q - 1`` (src/summit/sumstats/moments.py:24-139);
* per jackknife unit the sums ``m``, ``A_k``, ``(A^T A)``, ``(A^T L)`` and
  ``A^T y`` (src/summit/inference/h2core.py:453-543), replicates = full minus
  the deleted unit (``_stack_delete_replicates``, :612-627);
* the trace matrix ``T_kl = (A^T L)_kl N*^2 / (A_k A_l) + N*`` (the traces of
  the annotation kernels from null-subtracted LD scores,
  src/summit/utils.py:523-555), symmetrized pair by pair with the
  jackknife-variance-minimizing weight (src/summit/utils.py:1219-1445);
* the (K+1)-equation system ``[T, N*; N*, N*] sigma = [N* (A^T y)_k / A_k, N*]``
  per replicate (h2core.py:776-791), solved with LU, least squares when
  singular (h2core.py:545-566);
* per-bin heritability ``h2_k = sum_l (A^T A)_kl sigma_l / A_l``, the total
  ``sum_k sigma_k`` and enrichment ``(h2_k / h2) / (A_k / M)``
  (h2core.py:1182-1250);
* standard errors from the unequal-unit delete-one pseudovalues
  ``PV_u = (M theta - (M - m_u) theta_(-u)) / m_u`` with weights ``w_u = m_u/M``
  and variance ``sum_u w_u^2/(1 - w_u) (PV_u - sum w PV)^2``
  (src/summit/utils.py:1448-1690, exact delete-one branch).

Jackknife units are contiguous SNP blocks (equal SNP counts, or balanced on
the LD-score mass) or chromosomes (delete one chromosome), as
src/summit/inference/jackknife.py:17-214 builds them.
"""
from __future__ import annotations

import gzip
from dataclasses import dataclass

import numpy as np


# --------------------------------------------------------------------------- response
def effective_n_scale(nsamp: float, cov_rank: int = 0) -> float:
    """N* = N - q - 1 for q non-intercept covariates."""
    n = float(nsamp) - float(cov_rank) - 1.0
    if not (np.isfinite(n) and n > 0.0):
        raise ValueError(f"Invalid effective scale: nsamp={nsamp}, cov_rank={cov_rank}")
    return n


def score_z(beta, se, n_obs, nsamp: float, cov_rank: int = 0) -> np.ndarray:
    """Score statistic on the reference sample-size scale,
    ``z*_j = sqrt(N*) b_j / sqrt(b_j^2 + (n_j - q - 2) se_j^2)`` (NaN where undefined)."""
    beta = np.asarray(beta, dtype=np.float64)
    se = np.asarray(se, dtype=np.float64)
    n_obs = np.broadcast_to(np.asarray(n_obs, dtype=np.float64), beta.shape)
    n_star = effective_n_scale(nsamp, cov_rank)
    nu = n_obs - float(cov_rank) - 2.0
    den = beta * beta + nu * se * se
    ok = np.isfinite(beta) & np.isfinite(se) & (se > 0) & np.isfinite(n_obs) & (n_obs > 0) & (nu > 0) & (den > 0)
    z = np.full(beta.shape, np.nan)
    z[ok] = np.sqrt(n_star) * beta[ok] / np.sqrt(den[ok])
    return z


def score_noise_scale(n_obs, nsamp: float, cov_rank: int = 0, shape=None) -> np.ndarray:
    """Null variance of ``z*_j``: ``N* / (n_j - q - 1)`` (NaN where ``n_j - q - 1 <= 1``)."""
    n_obs = np.asarray(n_obs, dtype=np.float64)
    if shape is not None:
        n_obs = np.broadcast_to(n_obs, shape)
    n_star = effective_n_scale(nsamp, cov_rank)
    local = n_obs - float(cov_rank) - 1.0
    noise = np.full(n_obs.shape, np.nan)
    np.divide(n_star, local, out=noise, where=np.isfinite(local) & (local > 1.0))
    return noise


def h2_response(beta, se, n_obs, nsamp: float, cov_rank: int = 0) -> np.ndarray:
    """Per-SNP moment response from GWAS effect sizes and standard errors:
    z*^2 minus the excess of its null variance over one (NaN where undefined)."""
    z = score_z(beta, se, n_obs, nsamp, cov_rank)
    y = z * z - (score_noise_scale(n_obs, nsamp, cov_rank, shape=z.shape) - 1.0)
    y[~np.isfinite(y)] = np.nan
    return y


# --------------------------------------------------------------------------- jackknife units
@dataclass(frozen=True)
class JackknifeDesign:
    """Delete-one jackknife over contiguous SNP units (blocks or chromosomes):
    unit u holds SNPs [starts[u], ends[u]); replicate r deletes unit r."""
    mode: str
    starts: np.ndarray
    ends: np.ndarray

    @property
    def nunit(self) -> int:
        return int(self.starts.size)

    @property
    def nsnps(self) -> int:
        return int(self.ends[-1])

    def unit_id(self) -> np.ndarray:
        out = np.empty(self.nsnps, dtype=np.int64)
        for u, (s, e) in enumerate(zip(self.starts, self.ends)):
            out[s:e] = u
        return out

    @classmethod
    def blocks(cls, nsnps: int, nblocks: int = 200, balance: str = "snp", ldscores=None, annot=None,
               min_block_snps: int = 1) -> "JackknifeDesign":
        """Contiguous blocks: equal SNP counts (``floor(linspace(0, M, B+1))``), or
        cut at the weighted quantiles of a per-SNP mass (``ldscore``: the row sum
        of the LD scores; ``ldtrace``: (a_j . 1/A)(l_j . 1/A)) with at least
        min_block_snps per block (jackknife.py:110-211)."""
        M, B = int(nsnps), int(nblocks)
        if not (1 <= B <= M):
            raise ValueError(f"Need 1 <= nblocks <= nsnps, got B={B}, M={M}")
        balance = str(balance).strip().lower()
        if balance == "snp":
            return cls("block", *_equal_count_edges(M, B))
        if balance not in ("ldscore", "ldtrace"):
            raise ValueError(f"block_balance must be one of {{'snp','ldscore','ldtrace'}}, got {balance!r}")
        L = np.asarray(ldscores, dtype=np.float64)
        if balance == "ldscore":
            mass = L.sum(axis=1)
        else:
            A = np.asarray(annot, dtype=np.float64)
            Ak = A.sum(axis=0)
            inv = np.where(np.isfinite(Ak) & (Ak > 0), 1.0 / np.where(Ak > 0, Ak, 1.0), 0.0)
            mass = (A @ inv) * (L @ inv)
        mass = np.where(np.isfinite(mass) & (mass > 0), mass, 0.0)
        if min_block_snps <= 0 or min_block_snps * B > M:
            raise ValueError("min_block_snps must be >= 1 and min_block_snps * nblocks <= nsnps")
        total = float(mass.sum())
        if not (np.isfinite(total) and total > 0):
            return cls("block", *_equal_count_edges(M, B))
        cs = np.cumsum(mass)
        starts = np.empty(B, dtype=np.int64)
        ends = np.empty(B, dtype=np.int64)
        s = 0
        for b in range(B):
            starts[b] = s
            rem = B - b - 1
            if rem == 0:
                e = M
            else:
                e = int(np.searchsorted(cs, (b + 1) * total / B, side="left") + 1)
                e = min(max(e, s + min_block_snps), M - rem * min_block_snps)
            ends[b] = e
            s = e
        return cls("block", starts, ends)

    @classmethod
    def chromosomes(cls, chrom) -> "JackknifeDesign":
        """One unit per chromosome (SNPs sorted by chromosome), deleted one at a time."""
        c = np.asarray(chrom).ravel()
        if c.size > 1 and np.any(c[1:] < c[:-1]):
            raise ValueError("chr jackknife requires SNPs sorted by nondecreasing CHR.")
        labels = np.unique(c)
        if labels.size < 2:
            raise ValueError("delete-d must satisfy 1 <= d < #chromosomes")
        starts = np.searchsorted(c, labels, side="left").astype(np.int64)
        ends = np.searchsorted(c, labels, side="right").astype(np.int64)
        return cls("chr", starts, ends)


def _equal_count_edges(M: int, B: int):
    edges = np.floor(np.linspace(0, M, B + 1)).astype(np.int64)
    edges[-1] = M
    if np.any(edges[1:] <= edges[:-1]):
        raise RuntimeError("Equal-count partition produced an empty block.")
    return edges[:-1].copy(), edges[1:].copy()


def _replicates(unit: np.ndarray) -> np.ndarray:
    """Delete-one replicates then the full sum: (U + 1, ...)."""
    full = unit.sum(axis=0)
    return np.concatenate([full[None] - unit, full[None]], axis=0)


def _pseudovalues(reps: np.ndarray, sizes: np.ndarray) -> np.ndarray:
    """Unit pseudovalues (M theta - (M - m_u) theta_(-u)) / m_u from (U + 1, P) estimates."""
    M = float(sizes.sum())
    full = reps[-1]
    with np.errstate(invalid="ignore", divide="ignore"):
        return (M * full[None, :] - (M - sizes[:, None]) * reps[:-1]) / sizes[:, None]


def _jackknife_weights(sizes: np.ndarray, valid: np.ndarray):
    ww = sizes[valid] / float(sizes.sum())
    return ww / ww.sum()


def jackknife_se(reps: np.ndarray, sizes: np.ndarray) -> np.ndarray:
    """Delete-one SE of each column of (U + 1, P) replicate estimates (last row =
    full data) for units of sizes m_u (non-finite pseudovalues are left out)."""
    reps = np.asarray(reps, dtype=np.float64).reshape(reps.shape[0], -1)
    sizes = np.asarray(sizes, dtype=np.float64)
    M = float(sizes.sum())
    pv = _pseudovalues(reps, sizes)
    good_u = np.isfinite(sizes) & (sizes > 0) & (sizes < M)
    out = np.full(reps.shape[1], np.nan)
    for p in range(reps.shape[1]):
        if not np.isfinite(reps[-1, p]):
            continue
        valid = good_u & np.isfinite(pv[:, p])
        if valid.sum() <= 1:
            continue
        ww = _jackknife_weights(sizes, valid)
        vals = pv[valid, p]
        ctr = float(np.sum(ww * vals))
        term = (ww * ww / (1.0 - ww)) * (vals - ctr) ** 2
        if 1.0 / float(np.sum(ww * ww)) <= 1.0:
            continue
        out[p] = float(np.sqrt(np.sum(np.where(np.isfinite(term), term, 0.0))))
    return out


def _pair_moments(x_reps, y_reps, sizes):
    """Jackknife variances of x and y and their covariance from delete-one replicates."""
    reps = np.column_stack([x_reps, y_reps])
    M = float(sizes.sum())
    pv = _pseudovalues(reps, sizes)
    valid = np.isfinite(sizes) & (sizes > 0) & (sizes < M) & np.isfinite(pv).all(axis=1)
    if valid.sum() <= 1:
        return np.nan, np.nan, np.nan
    ww = _jackknife_weights(sizes, valid)
    if 1.0 / float(np.sum(ww * ww)) <= 1.0:
        return np.nan, np.nan, np.nan
    d = pv[valid] - (ww[:, None] * pv[valid]).sum(axis=0)[None, :]
    alpha = ww * ww / (1.0 - ww)
    return (float(np.sum(alpha * d[:, 0] ** 2)), float(np.sum(alpha * d[:, 1] ** 2)),
            float(np.sum(alpha * d[:, 0] * d[:, 1])))


def symmetrize_traces(trace: np.ndarray, sizes: np.ndarray) -> np.ndarray:
    """T_kl and T_lk of every replicate replaced by w T_kl + (1 - w) T_lk, w the
    weight minimizing the jackknife variance of the combination, clipped to [0, 1]
    (0.5 when the moments are undefined)."""
    out = np.array(trace, dtype=np.float64, copy=True)
    K = out.shape[1]
    for k in range(K):
        for l in range(k + 1, K):
            v1, v2, c12 = _pair_moments(trace[:, k, l], trace[:, l, k], sizes)
            w = 0.5
            den = v1 + v2 - 2.0 * c12
            if np.isfinite(den) and den > 0.0:
                w = float(np.clip((v2 - c12) / den, 0.0, 1.0)) if np.isfinite((v2 - c12) / den) else 0.5
            v = w * trace[:, k, l] + (1.0 - w) * trace[:, l, k]
            out[:, k, l] = v
            out[:, l, k] = v
    return out


def unit_sums(A, L, y, active, design: JackknifeDesign):
    """Per jackknife unit over the active SNPs: the count ``m``, ``A_k = sum a``,
    ``A^T A``, ``A^T L`` and ``A^T y`` (src/summit/inference/h2core.py:453-543,
    src/summit/inference/rgcore.py:528-541)."""
    U, K = design.nunit, A.shape[1]
    m_u = np.zeros(U)
    Ak_u = np.zeros((U, K))
    AA_u = np.zeros((U, K, K))
    AL_u = np.zeros((U, K, K))
    Ay_u = np.zeros((U, K))
    for u in range(U):
        s, e = int(design.starts[u]), int(design.ends[u])
        act = active[s:e]
        if not act.any():
            continue
        Au, Lu, yu = A[s:e][act], L[s:e][act], y[s:e][act]
        m_u[u] = float(Au.shape[0])
        Ak_u[u] = Au.sum(axis=0)
        AA_u[u] = Au.T @ Au
        AL_u[u] = Au.T @ Lu
        Ay_u[u] = Au.T @ yu
    return m_u, Ak_u, AA_u, AL_u, Ay_u


def _solve_each(lhs: np.ndarray, rhs: np.ndarray) -> np.ndarray:
    out = np.full(rhs.shape, np.nan)
    for i in range(lhs.shape[0]):
        A, b = lhs[i], rhs[i]
        if not (np.isfinite(A).all() and np.isfinite(b).all()):
            continue
        try:
            out[i] = np.linalg.solve(A, b)
        except np.linalg.LinAlgError:
            try:
                out[i] = np.linalg.lstsq(A, b, rcond=None)[0]
            except np.linalg.LinAlgError:
                continue
    return out


# --------------------------------------------------------------------------- the fit
@dataclass
class H2Fit:
    h2: float                 # total SNP heritability
    h2_se: float
    h2_cat: np.ndarray        # (K,) per-annotation heritability (overlap-aware)
    h2_cat_se: np.ndarray
    enrichment: np.ndarray    # (K,)
    enrichment_se: np.ndarray
    sigma: np.ndarray         # (K + 1,) per-bin variance components and the noise term
    sigma_se: np.ndarray
    n_scale: float
    nsnps: int
    reps_h2: np.ndarray       # (U + 1,) replicate totals, full estimate last
    sigma_reps: np.ndarray = None  # (U + 1, K + 1) replicate solutions, full estimate last
    unit_sizes: np.ndarray = None  # (U,) SNPs with a finite response per jackknife unit


def fit_h2(ldscores, annot, y, n_scale: float, design: JackknifeDesign) -> H2Fit:
    """Solve SUMMIT's univariate moment equations on SNPs with a finite response.

    ldscores: (M, K) null-subtracted LD scores of the K annotation bins (the
    engine's output); annot: (M, K) the annotation values of each SNP; y: (M,)
    the response (``h2_response``); n_scale: N*."""
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
    n = float(n_scale)
    if not (np.isfinite(n) and n > 0):
        raise RuntimeError(f"Invalid univariate n_scale={n_scale}")
    active = np.isfinite(y)
    U = design.nunit
    m_u, Ak_u, AA_u, AL_u, Ay_u = unit_sums(A, L, y, active, design)
    M_rep = _replicates(m_u)
    Ak_rep = _replicates(Ak_u)
    AA_rep = _replicates(AA_u)
    AL_rep = _replicates(AL_u)
    Ay_rep = _replicates(Ay_u)
    if np.any(~np.isfinite(Ak_rep[-1]) | (Ak_rep[-1] <= 0)):
        raise RuntimeError("One or more full-sample bins have non-positive total weight after matching.")
    denom = Ak_rep[:, :, None] * Ak_rep[-1][None, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        trace = AL_rep * n * n / denom + n
    trace = np.where((denom > 0) & np.isfinite(denom), trace, n)
    trace = symmetrize_traces(trace, m_u)
    lhs = np.full((U + 1, K + 1, K + 1), n)
    lhs[:, :K, :K] = trace
    rhs = np.full((U + 1, K + 1), n)
    with np.errstate(divide="ignore", invalid="ignore"):
        rhs[:, :K] = Ay_rep * n / Ak_rep
    rhs[:, :K][~np.isfinite(rhs[:, :K]) | (Ak_rep <= 0)] = np.nan
    sigma = _solve_each(lhs, rhs)
    sg = sigma[:, :K]
    h2_tot = np.nansum(sg, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        h2_cat = np.einsum("bkl,bl->bk", AA_rep, sg / Ak_rep)
        prop = Ak_rep / M_rep[:, None]
        enr = (h2_cat / h2_tot[:, None]) / prop
    enr[~np.isfinite(enr) | (h2_tot[:, None] <= 0)] = np.nan
    return H2Fit(h2=float(h2_tot[-1]), h2_se=float(jackknife_se(h2_tot[:, None], m_u)[0]),
                 h2_cat=h2_cat[-1], h2_cat_se=jackknife_se(h2_cat, m_u), enrichment=enr[-1],
                 enrichment_se=jackknife_se(enr, m_u), sigma=sigma[-1], sigma_se=jackknife_se(sigma, m_u),
                 n_scale=n, nsnps=int(active.sum()), reps_h2=h2_tot, sigma_reps=sigma, unit_sizes=m_u)


# --------------------------------------------------------------------------- the engine's files
def read_gw_ldscores(outpath: str):
    """(table columns, SNP ids, LD scores (M, K), bin names) of ``<outpath>.gw.ldscore.gz``
    as written by ``ldscore.write_gw_ldscore_table`` (CHR SNP BP, optional V,
    one column per bin, optional LD_REL_*)."""
    with gzip.open(f"{outpath}.gw.ldscore.gz", "rt") as f:
        header = f.readline().rstrip("\n").split("\t")
        rows = [line.rstrip("\n").split("\t") for line in f if line.strip()]
    skip = {"CHR", "SNP", "BP", "V"}
    bins = [c for c in header if c not in skip and not c.startswith("LD_REL_")]
    col = {c: i for i, c in enumerate(header)}
    L = np.array([[float(r[col[b]]) for b in bins] for r in rows], dtype=np.float64).reshape(len(rows), len(bins))
    chrom = np.array([int(r[col["CHR"]]) for r in rows], dtype=np.int64) if "CHR" in col else None
    snps = np.array([r[col["SNP"]] for r in rows]) if "SNP" in col else None
    return chrom, snps, L, bins


def read_gw_annotation_masses(outpath: str) -> np.ndarray:
    with open(f"{outpath}.gw.M") as f:
        return np.array([float(x) for x in f.readline().split()], dtype=np.float64)
