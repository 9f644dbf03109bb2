"""The headline shape itself (N=100,000 samples, M=500,000 SNPs, B=64 probes,
block-Markov genotypes, 0.5% missing): properties that hold at any size, since
no oracle finishes there.

* exact integer products: column sums of U = G.Z equal sum_i s_i z_ib with the
  SNP sums s_i of the statistics; the SNP-sum of W = G^T U equals
  sum_n U1_n U_bn with U1 = G.1; G.(Z1 - Z2) = G.Z1 - G.Z2;
* the FP64-semantics estimator: 5- and 6-limb LD scores agree (the fixed-point
  error of the 5-limb path, which is what separates it from FP64);
* phase 2's two missing-call methods (missing-flag MFMA plane, sums gathered
  over per-SNP lists in XCD-affine sample ranges) give bit-identical scores.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

N, M, B = 100_000, 500_000, 64


@pytest.fixture(scope="module")
def geno():
    from summit_amd import engine
    print(f"full-size genotypes N={N} M={M}", flush=True)
    g = engine.Genotype.synthetic(N, M, seed=303, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    yield g
    g.close()


def test_exact_products_checksums_at_full_size(geno):
    from summit_amd import engine
    dev = engine.default_device()
    gen = torch.Generator(device="cpu").manual_seed(11)
    z1 = (torch.randint(0, 2, (M, B), generator=gen, dtype=torch.int8) * 2 - 1).to(dev)
    z2 = (torch.randint(0, 2, (M, B), generator=gen, dtype=torch.int8) * 2 - 1).to(dev)
    u1 = engine.gz_i32(geno, z1)
    # column sums of U against the SNP sums of the statistics
    s = torch.from_numpy(geno.stats()["sum"]).to(dev)
    want = (s[:, None] * z1.to(torch.int64)).sum(dim=0)
    assert torch.equal(u1.to(torch.int64).sum(dim=1), want)
    # linearity in the probes
    u2 = engine.gz_i32(geno, z2)
    ud = engine.gz_i32(geno, (z1.to(torch.int16) - z2.to(torch.int16)).to(torch.int8))
    assert torch.equal(ud, u1 - u2)
    # SNP-sum of W = G^T U equals sum_n (G.1)_n U_bn
    w = engine.gtu_i64(geno, u1)
    ones = torch.ones((M, 1), dtype=torch.int8, device=dev)
    rs = engine.gz_i32(geno, ones)[0].to(torch.int64)
    assert torch.equal(w.sum(dim=0), (u1.to(torch.int64) * rs[None, :]).sum(dim=1))


def test_five_and_six_limbs_agree_at_full_size(geno):
    from summit_amd import ldscore
    r5 = ldscore.genomewide_ldscore(geno, B, seed=1303, limbs=5)
    r6 = ldscore.genomewide_ldscore(geno, B, seed=1303, limbs=6)
    assert r5.guard_counts == (0, 0) and r6.guard_counts == (0, 0)
    rel = np.max(np.abs(r5.ldscore - r6.ldscore) / np.maximum(np.abs(r6.ldscore), 1.0))
    print(f"max |L5 - L6| / max(|L6|, 1) = {rel:.3e}", flush=True)
    assert rel < 1e-10
    assert np.all(np.isfinite(r5.ldscore))


def test_phase2_missing_methods_identical_at_full_size(geno):
    from summit_amd import autotune, engine
    dev = engine.default_device()
    gen = torch.Generator(device="cpu").manual_seed(13)
    xz = torch.randn((B, N), generator=gen, dtype=torch.float64).to(dev)
    inv = engine.resvar_inv(geno, 0, M)
    key = autotune.p2_key(M, N, 1, B, 5)
    out = {}
    for method in (1, 2):
        autotune.set_value("p2_miss", key, method)
        ms = torch.zeros((M, 1), dtype=torch.float64, device=dev)
        engine.phase2_accum(geno, 0, M, inv, B, xz, ms, N)
        out[method] = ms.cpu().numpy()
    autotune.set_value("p2_miss", key, 0)
    assert np.array_equal(out[1], out[2])
    assert np.all(np.isfinite(out[1])) and np.abs(out[1]).max() > 0
