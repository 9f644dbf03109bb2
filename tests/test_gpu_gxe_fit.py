"""The GxE normal-equation inputs built on the GPU (summit_amd.gxe.summary_inputs)
against dense FP64 features of the same mean-imputed genotypes, and the
end-to-end fit of a simulated G + GxE + NxE phenotype."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import gxe_basis, gxe_dense_inputs, random_codes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _setup(oracle, m, n, seed, miss=0.01):
    from summit_amd import engine
    codes = random_codes(m, n, seed=seed, miss_rate=miss)
    packed = oracle.pack_codes(codes)
    g = engine.Genotype.from_bed_bytes(packed, n, m).prepare(impute="mean")
    X = oracle.read_block_standardized(packed, n, 0, m, None, 1, "mean", 0)  # (n, m), missing -> 0
    return g, X


@pytest.mark.parametrize("n_cov", [0, 2])
def test_summary_inputs_match_dense_features(oracle, n_cov):
    from summit_amd import gxe, gxe_fit
    m, n = 700, 520
    print(f"gxe fit inputs n_cov={n_cov}", flush=True)
    g, X = _setup(oracle, m, n, seed=97)
    rng = np.random.default_rng(11)
    env = rng.standard_normal(n) * 0.7 + 0.4
    C = gxe_basis(rng, n, n_cov) if n_cov else None
    pheno = rng.standard_normal(n) + X[:, :50] @ rng.normal(0, 0.1, 50)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.3) * 1.0])
    got = gxe.summary_inputs(g, env, pheno, C, annot, nvecs=16, seed=1303, step_size=250)
    ref, _ = gxe_dense_inputs(X, env, np.zeros((n, 0)) if C is None else C, pheno, annot)
    for key in ("score_x", "score_w", "norm_x", "norm_w", "diag_nxe_x", "diag_nxe_w"):
        np.testing.assert_allclose(got[key], ref[key], rtol=1e-9, atol=1e-12, err_msg=key)
    for key in ("residual_rank", "q_nxe", "q_residual", "trace_nxe", "trace_nxe_sq"):
        assert got[key] == pytest.approx(ref[key], rel=1e-10), key
    # with the same (stochastic) panels the assembled systems agree entry for entry
    for key in ("ld_xx", "ld_xw", "ld_wx", "ld_ww"):
        ref[key] = got[key]
    a = gxe_fit.assemble_normal_equations(**got)
    b = gxe_fit.assemble_normal_equations(**ref)
    np.testing.assert_allclose(a.matrix, b.matrix, rtol=1e-9)
    np.testing.assert_allclose(a.rhs, b.rhs, rtol=1e-9)
    np.testing.assert_allclose(a.traces, b.traces, rtol=1e-9)
    g.close()


def test_fit_recovers_simulated_components(oracle):
    """Independent SNPs (N=3,000, M=1,500), y = G b + (e.G) g + e.u + noise with
    variance 0.3 / 0.15 / 0.1 / 0.45, two covariates; LD panels from 64 probes;
    25-block jackknife."""
    from summit_amd import gxe, gxe_fit
    m, n = 1500, 3000
    print("gxe fit simulation", flush=True)
    g, X = _setup(oracle, m, n, seed=98, miss=0.005)
    rng = np.random.default_rng(12)
    env = rng.standard_normal(n)
    C = gxe_basis(rng, n, 2)
    hg, hx, hn = 0.3, 0.15, 0.1
    y = (X @ rng.normal(0, np.sqrt(hg / m), m) + (env[:, None] * X) @ rng.normal(0, np.sqrt(hx / m), m)
         + env * rng.normal(0, np.sqrt(hn), n) + rng.normal(0, np.sqrt(1 - hg - hx - hn), n))
    inp = gxe.summary_inputs(g, env, y, C, None, nvecs=64, seed=7, step_size=500)
    res = gxe_fit.fit(inp, blocks=np.repeat(np.arange(25), m // 25))
    truth = np.array([hg, hx, hn])
    se = res.proportion_standard_errors[:3]
    print("proportions", res.proportions, "se", res.proportion_standard_errors, flush=True)
    assert res.identifiable and np.all(np.isfinite(se)) and np.all(se > 0)
    assert np.all(np.abs(res.proportions[:3] - truth) < 4.0 * se + 0.03)
    g.close()
