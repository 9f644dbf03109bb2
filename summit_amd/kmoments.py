"""Unpartitioned GRM moments tr(K^p), p = 1..4, by sample-space probes
(the K-moment step of the reference's genome-wide estimator,
src/summit/ldscore/gw_ldscore.py:2387-2572, enabled by --write-kmoments).

K = P X diag(inv^2) X^T P / M is the working GRM of the LD-score run (P the
covariate projection when C/R are given).  Two applications of K to a panel of
q probes, Y1 = K Z and Y2 = K Y1, run on the GPU through the drop-in's device
GRM product (``gwldcore.grm_apply_dev``: exact int8-MFMA products, FP64 scales);
the panels stay in HBM between the two applications, and only the four per-probe
inner products come back to the host:
  z^T K z, ||K z||^2 = z^T K^2 z, (K z)^T K^2 z = z^T K^3 z, ||K^2 z||^2 = z^T K^4 z.
Their probe means estimate tr(K^p) (E[z z^T] = I).  With r = N_eff - 1 the rank of
the projected sample space, the summary measures K's departure from a rank-r
projection: for a spectrum lambda of r eigenvalues summing to r,
t0 = sum (lambda - 1)^2, t1 = sum lambda (lambda - 1)^2, t2 = sum lambda^2 (lambda - 1)^2,
and delta_reff = t0^2 / (t2 - 2 t1 + t0) = (sum (lambda - 1)^2)^2 / sum (lambda - 1)^4."""
from __future__ import annotations

import numpy as np
import torch

from . import engine, gwldcore

PROBE_SEED_XOR = 0xD1B54A32D192ED03


def probe_seed(root_seed: int) -> int:
    """The probe stream's seed: the run's root seed xor a fixed constant (gw_ldscore.py:2403)."""
    return int((int(root_seed) ^ PROBE_SEED_XOR) & 0xFFFFFFFFFFFFFFFF)


def sample_probes(n: int, q: int, rand_dist: str, seed: int) -> np.ndarray:
    """(n, q) FP64 probes from NumPy's default generator seeded with ``seed``, drawn
    in the reference's order so a run reproduces its probes (gw_ldscore.py:2404-2419):
    Rademacher from integers {0, 1} mapped to -1/+1, Gaussian, or Gaussian columns
    rescaled to norm sqrt(n) (spherical)."""
    rng = np.random.default_rng(seed)
    if rand_dist == "rademacher":
        z = rng.integers(0, 2, size=(n, q), dtype=np.int8).astype(np.float64)
        return np.asfortranarray(2.0 * z - 1.0)
    if rand_dist in ("gaussian", "normal", "spherical"):
        z = np.asfortranarray(rng.standard_normal(size=(n, q)))
        if rand_dist == "spherical":
            norms = np.linalg.norm(z, axis=0)
            good = norms > 0
            z[:, good] *= np.sqrt(n) / norms[good]
        return z
    raise ValueError(f"Unsupported rand_dist for K-moment estimation: {rand_dist}")


def trace_k2_from_ldscore(meansq_col: np.ndarray, nsnps: int, proj_rank: float) -> float:
    """tr(K^2) from an unpartitioned run's mean squared products: r^2 sum_j meansq_j / M^2
    (gw_ldscore.py:2195-2199; meansq before the null subtraction)."""
    return float(proj_rank) ** 2 * float(np.sum(meansq_col, dtype=np.float64)) / float(nsnps) ** 2


def _flag(err: float, ok: float, warn: float) -> str:
    return "OK" if err <= ok else ("WARN" if err <= warn else "BAD")


def summarize(k1: np.ndarray, k2: np.ndarray, k3: np.ndarray, k4: np.ndarray, trace_k2_used: float,
              proj_rank: float, nsnps: int, seed: int) -> dict:
    """The .gw.kmoments row from the per-probe inner products (gw_ldscore.py:2501-2537)."""
    q = int(np.size(k1))
    means = [float(np.mean(k)) for k in (k1, k2, k3, k4)]
    ses = [float(np.std(k, ddof=1) / np.sqrt(q)) if q > 1 else 0.0 for k in (k1, k2, k3, k4)]
    trK, trK2p, trK3, trK4 = means
    r = float(proj_rank)
    k2u = float(trace_k2_used)
    t0 = k2u - r
    t1 = trK3 - 2.0 * k2u + r
    t2 = trK4 - 2.0 * trK3 + k2u
    alpha = trK / r
    t0a = k2u - trK * trK / r
    t1a = trK3 - 2.0 * alpha * k2u + alpha * alpha * trK
    t2a = trK4 - 2.0 * alpha * trK3 + alpha * alpha * k2u
    alpha_err = abs(alpha - 1.0)
    k2_relerr = abs(trK2p - k2u) / max(abs(k2u), 1e-12)
    s4 = t2 - 2.0 * t1 + t0
    reff = (t0 * t0) / s4 if np.isfinite(s4) and s4 > 0.0 else float("nan")
    if np.isfinite(reff):
        spike = "LOW_EFFECTIVE_RANK" if reff < 50.0 else ("MODERATE" if reff < 200.0 else "DIFFUSE")
    else:
        spike = "NA"
    return {
        "nsnps": int(nsnps), "proj_rank": r, "num_probes": q, "probe_seed": int(seed),
        "trace_K_probe": trK, "trace_K_probe_se": ses[0], "trace_K_target_rank": r,
        "trace_K2_from_ldscore": k2u, "trace_K2_probe": trK2p, "trace_K2_probe_se": ses[1],
        "trace_K3_probe": trK3, "trace_K3_probe_se": ses[2], "trace_K4_probe": trK4, "trace_K4_probe_se": ses[3],
        "alpha_rank": 1.0, "alpha_probe": alpha, "t0_rank": t0, "t1_rank": t1, "t2_rank": t2,
        "t0_probealpha": t0a, "t1_probealpha": t1a, "t2_probealpha": t2a,
        "alpha_probe_err": alpha_err, "k2_probe_relerr": k2_relerr, "delta_reff_rank": reff,
        "alpha_flag": _flag(alpha_err, 5e-3, 1e-2), "mc_flag": _flag(k2_relerr, 5e-2, 1e-1), "spike_flag": spike,
    }


def probe_products(g: engine.Genotype, z: np.ndarray, inv_all, nsnps: int, step_size: int, C=None, R=None):
    """Per-probe z^T K z, z^T K^2 z, z^T K^3 z, z^T K^4 z for the columns of z (N, q)."""
    dev = engine.default_device()
    zq = torch.from_numpy(np.ascontiguousarray(np.asarray(z, dtype=np.float64).T)).to(dev)
    y1 = gwldcore.grm_apply_dev(g, zq, inv_all, nsnps, step_size, C, R)
    y2 = gwldcore.grm_apply_dev(g, y1, inv_all, nsnps, step_size, C, R)
    Z, Y1, Y2 = zq.cpu().numpy(), y1.cpu().numpy(), y2.cpu().numpy()
    return (np.einsum("qn,qn->q", Z, Y1), np.einsum("qn,qn->q", Y1, Y1), np.einsum("qn,qn->q", Y1, Y2),
            np.einsum("qn,qn->q", Y2, Y2))


def unpartitioned_kmoments(bed_prefix, fam_path, nsnps: int, step_size: int, inv_all, trace_k2_ldscore: float,
                           proj_rank: float, root_seed: int, *, rand_dist: str = "rademacher", num_probes: int = 128,
                           row_sel=None, ddof: int = 1, C=None, R=None, impute_mode: str = "mean",
                           impute_seed: int = 0, use_mailman: bool = False, outpath: str | None = None) -> dict:
    """The K-moment step of an unpartitioned run (gw_ldscore.py:2387-2572): probes, two
    GRM applications, traces and diagnostics; writes <outpath>.gw.kmoments when outpath
    is given.  use_mailman with HWE imputation selects the reference's Mailman GRM
    product, the same product here."""
    if trace_k2_ldscore is None:
        raise ValueError("trace_k2_from_ldscore must be provided for K-moment estimation.")
    if not np.isfinite(trace_k2_ldscore):
        raise ValueError(f"trace_k2_from_ldscore is not finite: {trace_k2_ldscore}")
    q = int(num_probes)
    if q <= 0:
        raise ValueError("num_probes must be positive.")
    if float(proj_rank) <= 0.0:
        raise ValueError(f"Projected rank must be positive; got {proj_rank}")
    mode = "hwe" if (use_mailman and impute_mode == "hwe") else impute_mode
    g, _, _ = gwldcore._geno(bed_prefix, fam_path, row_sel, ddof, mode, impute_seed)
    seed = probe_seed(root_seed)
    z = sample_probes(g.n, q, rand_dist, seed)
    if C is None or R is None:
        C = R = None
    k1, k2, k3, k4 = probe_products(g, z, inv_all, int(nsnps), int(step_size), C, R)
    row = summarize(k1, k2, k3, k4, trace_k2_ldscore, proj_rank, nsnps, seed)
    if outpath is not None:
        write_gw_kmoments(outpath, row)
    return row


def write_gw_kmoments(outpath, row: dict) -> str:
    """``<outpath>.gw.kmoments``: one tab-separated row with a header, %.10e
    (gw_ldscore.py:2539-2541)."""
    import pandas as pd

    path = f"{outpath}.gw.kmoments"
    pd.DataFrame([row]).to_csv(path, sep="\t", index=False, float_format="%.10e")
    return path
