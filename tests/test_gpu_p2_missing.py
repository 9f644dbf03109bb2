"""Phase 2 with missing calls (mean imputation): the missing-call sums gathered
over per-SNP missing lists and the missing-flag MFMA plane form the same exact
integers, so every phase-2 output is bit-identical between the two methods;
both are pinned to the oracle through the estimator tests."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _run(g, s, e, xz, V, B, limbs, method, mc=False, parts=0):
    from summit_amd import autotune, engine
    dev = engine.default_device()
    key = (e - s, g.n, B * ((V + 31) // 32 * 32), limbs)
    autotune.set_value("p2_miss", key, method)
    autotune.set_value("p2_miss_parts", key, parts)
    inv = engine.resvar_inv(g, s, e)
    meansq = torch.zeros((e - s, B), dtype=torch.float64, device=dev)
    agg = torch.zeros(B, dtype=torch.float64, device=dev) if mc else None
    per = torch.zeros((e - s, B), dtype=torch.float64, device=dev) if mc else None
    engine.phase2_accum(g, s, e, inv, V, xz, meansq, g.n, mc_aggregate_sumsq=agg, mc_per_snp_sumsq=per,
                        limbs=limbs, meansq_row0=s)
    autotune.set_value("p2_miss", key, 0)
    autotune.set_value("p2_miss_parts", key, 0)
    out = [meansq.cpu().numpy()]
    if mc:
        out += [agg.cpu().numpy(), per.cpu().numpy()]
    return out


@pytest.mark.parametrize("n,m,s,e,V,B,limbs,miss_rate", [
    (1000, 3000, 0, 3000, 64, 1, 5, 0.005),       # headline layout: one 64-column chunk
    (1037, 2900, 333, 2777, 40, 3, 5, 0.02),      # ragged samples / SNP range, three bins, padded columns
    (1100, 2000, 0, 2000, 100, 2, 6, 0.01),       # six limbs, columns past 64 per bin
    (513, 1500, 17, 1400, 7, 1, 5, 0.3),          # dense missingness, long per-SNP lists
])
def test_gathered_missing_sums_equal_missing_plane(n, m, s, e, V, B, limbs, miss_rate):
    from summit_amd import autotune, engine
    dev = engine.default_device()
    print(f"p2 missing methods N={n} M={m} [{s},{e}) V={V} B={B} limbs={limbs} miss={miss_rate}", flush=True)
    g = engine.Genotype.synthetic(n, m, seed=41, maf_lo=0.01, maf_hi=0.5, miss_rate=miss_rate, block_len=100,
                                  copy_prob=0.8)
    g.prepare(ddof=1, impute="mean")
    assert g.n_missing() > 0
    gen = torch.Generator(device="cpu").manual_seed(5)
    xz = torch.randn((B * V, n), generator=gen, dtype=torch.float64).to(dev)
    autotune.clear()
    plane = _run(g, s, e, xz, V, B, limbs, 1, mc=True)
    gathered = _run(g, s, e, xz, V, B, limbs, 2, mc=True)
    # per-SNP outputs bit-identical; the aggregate over SNPs is summed with
    # floating-point atomics (order not fixed), so it agrees to rounding
    assert np.array_equal(plane[0], gathered[0])
    assert np.array_equal(plane[2], gathered[2])
    np.testing.assert_allclose(plane[1], gathered[1], rtol=1e-13, atol=0)
    assert np.all(np.isfinite(plane[0])) and np.abs(plane[0]).max() > 0
    autotune.clear()
    g.close()


@pytest.mark.parametrize("parts", [1, 8, 32])
def test_gathered_missing_sums_any_sample_ranges(parts):
    """The sums are gathered per sample range (1, or a multiple of the XCD
    count): every range count gives the plane's results."""
    from summit_amd import autotune, engine
    dev = engine.default_device()
    n, m, V = 1500, 2500, 64
    print(f"p2 missing ranges={parts} N={n} M={m}", flush=True)
    g = engine.Genotype.synthetic(n, m, seed=45, maf_lo=0.01, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.8)
    g.prepare(ddof=1, impute="mean")
    gen = torch.Generator(device="cpu").manual_seed(7)
    xz = torch.randn((V, n), generator=gen, dtype=torch.float64).to(dev)
    autotune.clear()
    plane = _run(g, 0, m, xz, V, 1, 5, 1)
    gathered = _run(g, 0, m, xz, V, 1, 5, 2, parts=parts)
    assert np.array_equal(plane[0], gathered[0])
    autotune.clear()
    g.close()


def test_gathered_missing_sums_in_snp_slices():
    """More columns than the 4 GiB per-range sums buffer holds for all SNPs
    (16 ranges x 33,024 columns: slices of 1,024 SNPs): the gathered path runs
    in SNP slices and still equals the missing plane."""
    from summit_amd import autotune, engine
    dev = engine.default_device()
    n, m, V = 200, 5000, 33000
    print(f"p2 missing sliced N={n} M={m} V={V}", flush=True)
    g = engine.Genotype.synthetic(n, m, seed=43, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.5)
    g.prepare(ddof=1, impute="mean")
    gen = torch.Generator(device="cpu").manual_seed(6)
    xz = torch.randn((V, n), generator=gen, dtype=torch.float64).to(dev)
    autotune.clear()
    plane = _run(g, 0, m, xz, V, 1, 5, 1)
    gathered = _run(g, 0, m, xz, V, 1, 5, 2, parts=16)
    assert np.array_equal(plane[0], gathered[0])
    autotune.clear()
    g.close()


def test_gathered_missing_sums_one_hot_operand():
    """A one-hot operand (column c is 1 at sample c and 0 elsewhere) makes each
    sum at (SNP j, column c) the missing flag of sample c at j, so the gathered
    lists are compared call by call with the missing plane."""
    from summit_amd import autotune, engine
    dev = engine.default_device()
    n, m = 64, 700
    print(f"p2 missing one-hot N={n} M={m}", flush=True)
    g = engine.Genotype.synthetic(n, m, seed=47, maf_lo=0.05, maf_hi=0.5, miss_rate=0.05, block_len=100,
                                  copy_prob=0.5)
    g.prepare(ddof=1, impute="mean")
    xz = torch.eye(n, dtype=torch.float64, device=dev)
    autotune.clear()
    plane = _run(g, 0, m, xz, n, 1, 5, 1)
    gathered = _run(g, 0, m, xz, n, 1, 5, 2)
    assert np.array_equal(plane[0], gathered[0])
    autotune.clear()
    g.close()
