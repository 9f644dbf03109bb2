"""Shared input generators for the tests (seeded, small)."""
from __future__ import annotations

import numpy as np


def random_codes(m: int, n: int, seed: int, miss_rate: float = 0.02, maf_lo: float = 0.05,
                 maf_hi: float = 0.5, monomorphic: int = 0, all_missing: int = 0) -> np.ndarray:
    """SNP-major 2-bit codes (m, n): HWE genotypes, missing calls, optional
    monomorphic and all-missing SNPs (the edge cases the reference tests use)."""
    rng = np.random.default_rng(seed)
    maf = rng.uniform(maf_lo, maf_hi, size=m)
    g = rng.binomial(2, maf[:, None], size=(m, n))
    codes = np.array([0, 2, 3], dtype=np.uint8)[g]
    codes[rng.random((m, n)) < miss_rate] = 1
    for j in range(min(monomorphic, m)):
        codes[j, :] = 0 if j % 2 == 0 else 3
    for j in range(min(all_missing, m)):
        codes[m - 1 - j, :] = 1
    return codes


def assert_ld_close(ld, ref, null, rtol: float = 1e-10):
    """LD scores L = meansq/V - null_k against the oracle: |dL| <= rtol (|L| + null_k)
    per element, i.e. rtol relative to the accumulated statistic meansq/V = L + null_k
    that the estimator computes (the null subtraction M_k/N_denom can cancel it to
    near zero, where a bare relative bar on L would measure the cancellation)."""
    ld = np.asarray(ld, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    scale = np.abs(ref) + np.abs(np.asarray(null, dtype=np.float64))[None, :]
    err = np.abs(ld - ref)
    bad = err > rtol * scale
    assert not np.any(bad), (f"{int(bad.sum())} LD scores beyond {rtol:g} relative; worst "
                             f"{float(np.max(err / np.maximum(scale, 1e-300))):.3g}")


def gxe_basis(rng, n, p):
    """[1/sqrt(N), C] with C orthonormal and centred."""
    C = rng.standard_normal((n, p))
    C -= C.mean(axis=0)
    C, _ = np.linalg.qr(C)
    return C


def gxe_features(G, e, C):
    n = G.shape[0]
    Q = np.column_stack([np.full(n, 1.0 / np.sqrt(n)), C])
    P = np.eye(n) - Q @ Q.T
    r = n - Q.shape[1]
    fx = P @ G
    fw = P @ (e[:, None] * G)
    fx *= np.sqrt(r / np.sum(fx * fx, axis=0))
    fw *= np.sqrt(r / np.sum(fw * fw, axis=0))
    return P, r, fx, fw


def gxe_dense_inputs(G, e, C, y_raw, annot):
    """Exact summary arrays from dense features (the panels without probes)."""
    P, r, fx, fw = gxe_features(G, e, C)
    y = P @ y_raw
    y *= np.sqrt(r / (y @ y))
    d = e * e
    ld = {}
    for key, (lf, sf) in {"xx": (fx, fx), "xw": (fx, fw), "wx": (fw, fx), "ww": (fw, fw)}.items():
        ld[key] = ((lf.T @ sf / r) ** 2) @ annot
    from summit_amd import gxe_fit
    t1, t2 = gxe_fit.nxe_traces(e, C)
    return {"annotations": annot, "score_x": fx.T @ y / np.sqrt(r), "score_w": fw.T @ y / np.sqrt(r),
            "ld_xx": ld["xx"], "ld_xw": ld["xw"], "ld_wx": ld["wx"], "ld_ww": ld["ww"],
            "norm_x": np.sum(fx * fx, axis=0) / r, "norm_w": np.sum(fw * fw, axis=0) / r,
            "diag_nxe_x": np.sum(d[:, None] * fx * fx, axis=0) / r,
            "diag_nxe_w": np.sum(d[:, None] * fw * fw, axis=0) / r,
            "residual_rank": r, "q_nxe": float(np.sum((e * y) ** 2)), "q_residual": float(y @ y),
            "trace_nxe": t1, "trace_nxe_sq": t2}, (P, r, fx, fw, y)
