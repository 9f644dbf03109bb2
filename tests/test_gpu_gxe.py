"""SUMMIT-GxE directional scores on the GPU against the oracle's restatement of the
reference's Python backend."""
from __future__ import annotations

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_gxe_probes_match_numpy(oracle):
    from summit_amd import gxe
    z = gxe.probes_device(1505, 0, 2300, 1000, [5, 6, 7]).cpu().numpy()
    ref = np.concatenate([oracle.gxe_probe_block(1505, s, [5, 6, 7], min(2300, s + 1000) - s)
                          for s in range(0, 2300, 1000)])
    np.testing.assert_array_equal(z.astype(np.float64), ref)


@pytest.mark.parametrize("with_cov", [False, True])
def test_gxe_scores_match_oracle(oracle, with_cov):
    from summit_amd import engine, gxe
    m, n, nvecs, step = 600, 420, 12, 250
    codes = random_codes(m, n, seed=91, miss_rate=0.01)
    rng = np.random.default_rng(5)
    env = rng.standard_normal(n)
    env = (env - env.mean()) / env.std()
    C = None
    if with_cov:
        raw = np.column_stack([rng.standard_normal(n), rng.binomial(1, 0.3, n)]).astype(float)
        raw -= raw.mean(axis=0)
        C, _ = np.linalg.qr(raw)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.2).astype(float)])
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(impute="mean")
    got = gxe.gxe_ldscore(g, env, C, annot, nvecs=nvecs, seed=1303, step_size=step, vtiles=[8, 4])
    ref = oracle.gxe_scores(oracle.pack_codes(codes), n, m, env, C, annot, nvecs=nvecs, seed=1303,
                            step_size=step, vtiles=[8, 4])
    np.testing.assert_allclose(got["inv_x"], ref["inv_x"], rtol=1e-12)
    np.testing.assert_allclose(got["inv_w"], ref["inv_w"], rtol=1e-12)
    for k in ("xx", "xw", "wx", "ww"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-10, err_msg=k)
    g.close()


@pytest.mark.parametrize("dist,fp32", [("gaussian", True), ("gaussian", False), ("spherical", False)])
def test_gxe_real_valued_probes_match_oracle(oracle, dist, fp32):
    """Gaussian (float32-rounded as a float32 run stores them, or FP64) and spherical
    Philox probes drawn on the host, through the FP64-probe phase 1."""
    from summit_amd import engine, gxe
    m, n, nvecs, step = 500, 300, 10, 170
    print(f"gxe real probes {dist} fp32={fp32}", flush=True)
    codes = random_codes(m, n, seed=92, miss_rate=0.01)
    rng = np.random.default_rng(6)
    env = (rng.random(n) < 0.3).astype(float)
    raw = rng.standard_normal((n, 2))
    raw -= raw.mean(axis=0)
    C, _ = np.linalg.qr(raw)
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.3).astype(float)])
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(impute="mean")
    got = gxe.gxe_ldscore(g, env, C, annot, nvecs=nvecs, seed=77, step_size=step, vtiles=[6, 4], rand_dist=dist,
                          probe_fp32=fp32)
    ref = oracle.gxe_scores(oracle.pack_codes(codes), n, m, env, C, annot, nvecs=nvecs, seed=77, step_size=step,
                            vtiles=[6, 4], rand_dist=dist, probe_fp32=fp32)
    for k in ("xx", "xw", "wx", "ww"):
        np.testing.assert_allclose(got[k], ref[k], rtol=1e-10, err_msg=k)
    g.close()


@pytest.mark.parametrize("per_env_cov", [False, True])
def test_gxe_multi_environment_matches_oracle(oracle, per_env_cov):
    """Several environments served by one set of traversals (the reference's
    MultiEnvironmentKernel) against the oracle run environment by environment."""
    from summit_amd import engine, gxe
    m, n, nvecs, step = 520, 380, 10, 200
    codes = random_codes(m, n, seed=93, miss_rate=0.01)
    rng = np.random.default_rng(8)
    envs = np.column_stack([rng.standard_normal(n), rng.standard_normal(n), (rng.random(n) < 0.3).astype(float)])
    envs[:, :2] = (envs[:, :2] - envs[:, :2].mean(axis=0)) / envs[:, :2].std(axis=0)
    raw = np.column_stack([rng.standard_normal(n), rng.binomial(1, 0.4, n)]).astype(float)
    raw -= raw.mean(axis=0)
    C, _ = np.linalg.qr(raw)
    Cs = [C, None, C[:, :1]] if per_env_cov else C
    annot = np.column_stack([np.ones(m), (rng.random(m) < 0.25).astype(float)])
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(impute="mean")
    got = gxe.gxe_ldscore_multi(g, envs, Cs, annot, nvecs=nvecs, seed=77, step_size=step, vtiles=[6, 4])
    for k in range(envs.shape[1]):
        Ck = Cs[k] if per_env_cov else C
        ref = oracle.gxe_scores(oracle.pack_codes(codes), n, m, envs[:, k], Ck, annot, nvecs=nvecs, seed=77,
                                step_size=step, vtiles=[6, 4])
        np.testing.assert_allclose(got[k]["inv_x"], ref["inv_x"], rtol=1e-12)
        np.testing.assert_allclose(got[k]["inv_w"], ref["inv_w"], rtol=1e-12)
        for name in ("xx", "xw", "wx", "ww"):
            np.testing.assert_allclose(got[k][name], ref[name], rtol=1e-10, err_msg=f"env {k} {name}")
    g.close()


@pytest.mark.parametrize("n_w,impute", [(1, "mean"), (3, "mean"), (17, "mean"), (40, "mean"), (5, "hwe")])
def test_weighted_sumsq_matches_dense(oracle, n_w, impute):
    """sum_n w_n x_jn^2 for several weight vectors in one call (plane sums on int8
    MFMA, two column groups at 40 weights) against the oracle's standardized matrix;
    a SNP range that starts mid-matrix; signed weights; HWE-completed codes."""
    from summit_amd import engine, gxe
    m, n = 300, 1037
    codes = random_codes(m, n, seed=41, miss_rate=0.03, monomorphic=1)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(impute=impute, impute_seed=5)
    rng = np.random.default_rng(n_w)
    W = rng.standard_normal((n_w, n)) ** 2
    W[-1] -= 0.5  # a weight vector with both signs
    X = oracle.read_block_standardized(oracle.pack_codes(codes), n, 0, m, None, 1, impute, 5)
    got = gxe.weighted_sumsq(g, torch.as_tensor(W, device=engine.default_device()), 20, 290).cpu().numpy()
    ref = W @ (X[:, 20:290] ** 2)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-9)
    g.close()


def test_phase1_int8_probes_must_be_signs(oracle):
    """gpe_phase1_xz_z8 takes caller-provided int8 probes; a value outside
    {-1, 0, +1} would not fit its missing-call records and is refused."""
    from summit_amd import _native, engine
    m, n, V = 300, 200, 8
    codes = random_codes(m, n, seed=61, miss_rate=0.02)
    g = engine.Genotype.from_bed_bytes(oracle.pack_codes(codes), n, m).prepare(ddof=1, impute="mean")
    dev = engine.default_device()
    annot = torch.ones((m, 1), dtype=torch.float64, device=dev)
    inv = engine.resvar_inv(g, 0, m)
    z = torch.from_numpy(np.random.default_rng(1).choice([-1, 0, 1], size=(m, V)).astype(np.int8)).to(dev)
    xz = torch.zeros((V, n), dtype=torch.float64, device=dev)
    engine.call("gpe_phase1_xz_z8", g.handle, 0, m, engine._ptr(annot), 1, engine._ptr(inv), engine._ptr(z), V, V,
                engine._ptr(xz), n, None, 5, engine._stream())
    z[123, 5] = 2
    with pytest.raises(_native.GpeError, match="int8 probes must be -1, 0 or \\+1"):
        engine.call("gpe_phase1_xz_z8", g.handle, 0, m, engine._ptr(annot), 1, engine._ptr(inv), engine._ptr(z), V,
                    V, engine._ptr(xz), n, None, 5, engine._stream())
    g.close()
