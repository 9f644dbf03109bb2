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
