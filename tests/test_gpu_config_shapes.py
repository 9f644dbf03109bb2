"""The benchmark configurations at their stated shapes, checked through
properties that hold at any size (the dense oracle cannot run there):

* partitioned LD scores, N=50,000, M=200,000, 8 random 10% annotations,
  B=100: the 5-limb and 6-limb fixed-point paths agree to 1e-10 per score
  (relative to the accumulated statistic), and the errors against the exact
  partitioned LD scores (full-window int8 Gram) behave as Monte Carlo noise:
  standardized by the estimator's own Monte Carlo standard errors they have
  variance ~1 and at most 0.1% beyond 5, and no bin's mean error exceeds 3
  typical standard errors;
* GxE, N=100,000, M=500,000, B=64, K=8 environments (6 N(0,1), 2 Bernoulli(0.3)),
  Rademacher and Gaussian float32 probes: every environment served by the
  batched traversals equals its own one-environment run; and at a size the
  oracle reaches, K=8 environments on block-Markov genotypes match it;
* windowed LD, N=50,000, M=300,000 on 22 chromosomes with Poisson positions and
  +-1 Mb windows: the one-launch result equals the per-chromosome launches
  (to FP64 round-off: tile pairs add into each SNP's score by FP64 atomics,
  in an order that varies from launch to launch),
  the windows span the stated 200-5,000 SNPs, and SNPs spread over the
  chromosomes match a dense computation from their windows' standardized
  genotypes (formed by the engine's own X.c product with c = I)."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _markov(engine, N, M, seed):
    g = engine.Genotype.synthetic(N, M, seed=seed, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    return g.prepare(ddof=1, impute="mean")


def test_partitioned_full_shape():
    from summit_amd import cv, engine, ldscore
    from tests.helpers import assert_ld_close
    N, M, B, K = 50_000, 200_000, 100, 8
    print(f"partitioned N={N} M={M} B={B} K={K}", flush=True)
    g = _markov(engine, N, M, 202)
    annot = (np.random.default_rng(1202).random((M, K)) < 0.1).astype(np.float64)
    r5 = ldscore.genomewide_ldscore(g, B, seed=1202, annot=annot, limbs=5, mc=True)
    r6 = ldscore.genomewide_ldscore(g, B, seed=1202, annot=annot, limbs=6, mc=True)
    assert_ld_close(r5.ldscore, r6.ldscore, annot.sum(axis=0) / (N - 1.0))
    exact = cv.exact_ldscore(g, annot)
    g.close()
    var = np.maximum(r5.mc_var, 0.0)
    err = r5.ldscore - exact
    z = err / np.maximum(np.sqrt(var), 1e-300)
    frac5 = float(np.mean(np.abs(z) <= 5.0))
    # no systematic error: per bin, the mean error over SNPs is within 3 typical standard
    # errors, a bound that holds however the SNPs' errors correlate (they share the
    # probes); the mean of z itself is not a bias test, since each SE is estimated from
    # the same probes as its score and the per-probe values are skewed
    bias = np.abs(err.mean(axis=0)) / np.sqrt(var.mean(axis=0))
    print(f"z: mean {z.mean():.4f} var {z.var():.4f} within 5 SE {frac5:.6f}; "
          f"|mean error| / typical SE per bin {np.round(bias, 4).tolist()}", flush=True)
    assert frac5 >= 0.999
    assert np.all(bias < 3.0)
    assert 0.8 < float(z.var()) < 1.25


def _envs(N, K, seed):
    rng = np.random.default_rng(seed)
    envs = np.empty((N, K))
    for k in range(K - 2):
        e = rng.standard_normal(N)
        envs[:, k] = (e - e.mean()) / e.std()
    envs[:, K - 2:] = (rng.random((N, 2)) < 0.3).astype(np.float64)
    return envs


@pytest.mark.parametrize("dist,fp32", [("rademacher", False), ("gaussian", True)])
def test_gxe_full_shape_batched_equals_single(dist, fp32):
    from summit_amd import engine, gxe
    N, M, B, K = 100_000, 500_000, 64, 8
    print(f"gxe N={N} M={M} B={B} K={K} {dist} fp32={fp32}", flush=True)
    g = _markov(engine, N, M, 303)
    envs = _envs(N, K, 1303)
    multi = gxe.gxe_ldscore_multi(g, envs, None, nvecs=B, seed=1303, rand_dist=dist, probe_fp32=fp32)
    for k in (0, K - 1):  # a continuous and a binary environment
        one = gxe.gxe_ldscore(g, envs[:, k], None, nvecs=B, seed=1303, rand_dist=dist, probe_fp32=fp32)
        for name in ("xx", "xw", "wx", "ww"):
            np.testing.assert_allclose(multi[k][name], one[name], rtol=1e-12, atol=1e-12, err_msg=f"{k} {name}")
    # XX does not depend on the environment: every environment's XX is the same estimate
    for k in range(1, K):
        np.testing.assert_allclose(multi[k]["xx"], multi[0]["xx"], rtol=1e-12, atol=1e-12)
    for k in range(K):
        for name in ("xx", "xw", "wx", "ww"):
            assert np.isfinite(multi[k][name]).all()
    g.close()


def test_gxe_eight_environments_block_markov_vs_oracle(oracle):
    from summit_amd import engine, gxe
    N, M, B, K = 1500, 2000, 16, 8
    g = _markov(engine, N, M, 7)
    packed = g.download("snp_raw")
    envs = _envs(N, K, 8)
    got = gxe.gxe_ldscore_multi(g, envs, None, nvecs=B, seed=99, step_size=300, vtiles=[10, 6])
    g.close()
    for k in (0, 3, K - 2, K - 1):
        ref = oracle.gxe_scores(packed, N, M, envs[:, k], None, None, nvecs=B, seed=99, step_size=300,
                                vtiles=[10, 6])
        for name in ("xx", "xw", "wx", "ww"):
            np.testing.assert_allclose(got[k][name], ref[name], rtol=1e-10, err_msg=f"{k} {name}")


def test_windowed_full_shape(oracle):
    from summit_amd import engine, winldcore
    N, M, n_chr = 50_000, 300_000, 22
    print(f"windowed N={N} M={M} chromosomes={n_chr}", flush=True)
    g = _markov(engine, N, M, 404)
    rng = np.random.default_rng(1404)
    bounds = np.linspace(0, M, n_chr + 1).astype(np.int64)
    chrs, sizes = [], []
    for c in range(n_chr):
        s0, s1 = int(bounds[c]), int(bounds[c + 1])
        bp = np.cumsum(rng.exponential(3000.0, s1 - s0)).astype(np.int64) + 1
        lo = np.searchsorted(bp, bp - 1_000_000, side="left")
        hi = np.searchsorted(bp, bp + 1_000_000, side="right")
        sizes.append(hi - lo)
        chrs.append((s0, s1, bp))
    sizes = np.concatenate(sizes)
    assert sizes.min() >= 200 and sizes.max() <= 5000
    allc = winldcore.windowed_ld_multi(g, chrs, ld_wind_kb=1000.0)
    for c in (0, 11, 21):
        s0, s1, bp = chrs[c]
        one = winldcore.windowed_ld(g, s0, s1, bp=bp, ld_wind_kb=1000.0)
        print(f"chromosome {c}: one launch vs per chromosome, max |diff| {np.max(np.abs(allc[s0:s1] - one)):.3e}",
              flush=True)
        np.testing.assert_allclose(allc[s0:s1], one, rtol=1e-12, atol=1e-12)
    # dense spot checks: SNPs at the start, middle and end of a few chromosomes
    dev = engine.default_device()
    for c, pos in ((0, 0), (5, None), (21, -1)):
        s0, s1, bp = chrs[c]
        j = {0: 0, -1: s1 - s0 - 1}.get(pos, (s1 - s0) // 2)
        lo = int(np.searchsorted(bp, bp[j] - 1_000_000, side="left"))
        hi = int(np.searchsorted(bp, bp[j] + 1_000_000, side="right"))
        L = hi - lo
        eye = torch.eye(L, dtype=torch.float64, device=dev)
        X = engine.xc(g, eye, s0 + lo, s0 + hi, limbs=6).cpu().numpy().T  # (N, L) standardized columns
        Xr = oracle.restandardize_cols(X)
        r = Xr.T @ Xr[:, j - lo] / N
        r2 = r * r
        want = float(np.sum(r2 - (1.0 - r2) / (N - 2)))
        got = float(allc[s0 + j, 0])
        print(f"chromosome {c} SNP {j}: window {L}, dense {want:.10f}, engine {got:.10f}", flush=True)
        assert got == pytest.approx(want, rel=1e-9, abs=1e-9)
    g.close()
