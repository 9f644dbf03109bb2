"""The GxE variance-component normal equations (summit_amd.gxe_fit) against the
kernel matrices they summarize, built explicitly on small inputs, and against
simulated phenotypes with known G, GxE and NxE variance (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from summit_amd import gxe_fit
from tests.helpers import gxe_basis as _basis, gxe_dense_inputs as _inputs


def _small(seed=1, n=150, m=40, p=2):
    rng = np.random.default_rng(seed)
    G = rng.binomial(2, rng.uniform(0.1, 0.5, m), size=(n, m)).astype(float)
    G = (G - G.mean(0)) / G.std(0)
    e = rng.standard_normal(n) * 0.8 + 0.3
    C = _basis(rng, n, p)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.4) * rng.uniform(0.5, 2.0, m)])
    annot[0, 1] = 1.0
    return G, e, C, rng.standard_normal(n), annot


def test_assembly_is_the_kernel_gram_matrix():
    """tr(K_c K_d), y'K_c y and tr K_c of the explicit kernels K_G,a, K_GxE,a,
    P diag(e^2) P and P equal the assembled system built from per-SNP sums."""
    G, e, C, y0, annot = _small()
    inp, (P, r, fx, fw, y) = _inputs(G, e, C, y0, annot)
    eq = gxe_fit.assemble_normal_equations(**inp)
    mass = annot.sum(axis=0)
    kern = [(fx * annot[:, a]) @ fx.T / mass[a] for a in range(2)]
    kern += [(fw * annot[:, a]) @ fw.T / mass[a] for a in range(2)]
    kern += [P @ np.diag(e * e) @ P, P]
    lhs = np.array([[np.sum(Kc * Kd) for Kd in kern] for Kc in kern])
    np.testing.assert_allclose(eq.matrix, lhs, rtol=1e-10)
    np.testing.assert_allclose(eq.rhs, [y @ K @ y for K in kern], rtol=1e-10)
    np.testing.assert_allclose(eq.traces, [np.trace(K) for K in kern], rtol=1e-10)
    assert eq.component_names == ("G:bin0", "G:bin1", "GxE:bin0", "GxE:bin1", "NxE", "residual")


def test_nxe_traces_brute_force():
    rng = np.random.default_rng(2)
    n = 90
    e = rng.standard_normal(n)
    C = _basis(rng, n, 3)
    Q = np.column_stack([np.full(n, 1.0 / np.sqrt(n)), C])
    K = (np.eye(n) - Q @ Q.T) @ np.diag(e * e) @ (np.eye(n) - Q @ Q.T)
    t1, t2 = gxe_fit.nxe_traces(e, C)
    assert t1 == pytest.approx(np.trace(K), rel=1e-12)
    assert t2 == pytest.approx(np.sum(K * K), rel=1e-12)
    # no covariates: only the intercept
    K0 = (np.eye(n) - 1.0 / n) @ np.diag(e * e) @ (np.eye(n) - 1.0 / n)
    assert gxe_fit.nxe_traces(e)[1] == pytest.approx(np.sum(K0 * K0), rel=1e-12)


def test_solve_matches_the_full_system():
    """Eliminating the residual row gives the solution of the full system, and
    the residual coefficient closes the y'y = r equation."""
    G, e, C, y0, annot = _small(seed=3)
    inp, _ = _inputs(G, e, C, y0, annot)
    eq = gxe_fit.assemble_normal_equations(**inp)
    res = gxe_fit.solve_normal_equations(eq, allow_ill_conditioned=True)
    np.testing.assert_allclose(res.coefficients, np.linalg.solve(eq.matrix, eq.rhs), rtol=1e-7, atol=1e-9)
    assert res.coefficients[-1] == pytest.approx(1.0 - eq.traces[:-1] @ res.coefficients[:-1] / eq.traces[-1])
    assert res.relative_residual < 1e-10
    assert res.rank == eq.matrix.shape[0]
    np.testing.assert_allclose(res.proportions.sum(), 1.0, rtol=1e-12)


def test_solve_rejects_inconsistent_systems():
    G, e, C, y0, annot = _small(seed=4)
    inp, _ = _inputs(G, e, C, y0, annot)
    eq = gxe_fit.assemble_normal_equations(**inp)
    A = eq.matrix.copy()
    A[0, 1] *= 1.01  # asymmetric
    with pytest.raises(ValueError, match="symmetric"):
        gxe_fit.solve_normal_equations(gxe_fit.GxENormalEquations(A, eq.rhs, eq.traces, eq.component_names))
    B = eq.matrix.copy()
    B[0, 1] = B[1, 0] = 2.0 * np.sqrt(B[0, 0] * B[1, 1])  # not a Gram matrix
    with pytest.raises(ValueError, match="Cauchy-Schwarz"):
        gxe_fit.solve_normal_equations(gxe_fit.GxENormalEquations(B, eq.rhs, eq.traces, eq.component_names))
    bad = dict(inp, q_residual=inp["q_residual"] * 1.1)
    with pytest.raises(ValueError, match="q_residual"):
        gxe_fit.assemble_normal_equations(**bad)
    with pytest.raises(ValueError, match="non-negative"):
        gxe_fit.assemble_normal_equations(**dict(inp, ld_xx=-inp["ld_xx"]))
    # a constant environment makes NxE the residual kernel up to a factor
    e1 = np.full_like(e, 1.7)
    inp1, _ = _inputs(G, e1, C, y0, annot)
    with pytest.raises(ValueError, match="identifiable"):
        gxe_fit.solve_normal_equations(gxe_fit.assemble_normal_equations(**inp1))


def test_deleted_blocks_equal_reassembly_without_the_block():
    """A deleted system is the full assembly on the SNPs outside the block, the
    panels (full-genome sums over sources) left as they are."""
    G, e, C, y0, annot = _small(seed=5, m=60)
    inp, _ = _inputs(G, e, C, y0, annot)
    blocks = np.repeat(np.arange(6), 10)
    dels = gxe_fit.deleted_block_equations(inp, blocks)
    assert len(dels) == 6
    for jb in (0, 3, 5):
        keep = blocks != jb
        sub = {k: (v[keep] if isinstance(v, np.ndarray) and v.shape[:1] == (60,) else v) for k, v in inp.items()}
        ref = gxe_fit.assemble_normal_equations(**sub)
        np.testing.assert_allclose(dels[jb].matrix, ref.matrix, rtol=1e-12)
        np.testing.assert_allclose(dels[jb].rhs, ref.rhs, rtol=1e-12)
        np.testing.assert_allclose(dels[jb].traces, ref.traces, rtol=1e-12)
    with pytest.raises(ValueError):
        gxe_fit.deleted_block_equations(inp, np.zeros(60, dtype=int))


def test_simulated_components_recovered():
    """y = G b + (e.G) g + e.u + noise with variances 0.3 / 0.15 / 0.1 / rest:
    the fitted proportions come back within 4 jackknife SEs."""
    rng = np.random.default_rng(6)
    n, m = 3000, 500
    G = rng.binomial(2, rng.uniform(0.05, 0.5, m), size=(n, m)).astype(float)
    G = (G - G.mean(0)) / G.std(0)
    e = rng.standard_normal(n)
    C = _basis(rng, n, 2)
    hg, hx, hn = 0.3, 0.15, 0.1
    y = (G @ rng.normal(0, np.sqrt(hg / m), m) + (e[:, None] * G) @ rng.normal(0, np.sqrt(hx / m), m)
         + e * rng.normal(0, np.sqrt(hn), n) + rng.normal(0, np.sqrt(1 - hg - hx - hn), n))
    inp, _ = _inputs(G, e, C, y, np.ones((m, 1)))
    res = gxe_fit.fit(inp, blocks=np.repeat(np.arange(25), m // 25))
    truth = np.array([hg, hx, hn, 1 - hg - hx - hn])
    se = res.proportion_standard_errors
    assert res.identifiable and np.all(np.isfinite(se)) and np.all(se[:3] > 0)
    assert np.all(np.abs(res.proportions - truth)[:3] < 4.0 * se[:3] + 0.02), (res.proportions, se)
    assert res.proportions[0] > 0.15 and res.proportions[1] > 0.05
