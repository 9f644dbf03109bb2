"""The schedule autotuner on the GPU: it times candidates, keeps the fastest in
the library's table and cache, and a tuned schedule gives the same phase-1
products as the heuristic one."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def test_tune_phase1_caches_and_keeps_results():
    from summit_amd import autotune, engine
    dev = engine.default_device()
    N, M, V = 2000, 5000, 40
    print(f"autotune phase 1 N={N} M={M} V={V}", flush=True)
    g = engine.Genotype.synthetic(N, M, seed=9, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100, copy_prob=0.5)
    g.prepare(ddof=1, impute="mean")
    annot = torch.ones((M, 1), dtype=torch.float64, device=dev)
    inv = engine.resvar_inv(g, 0, M)
    key = autotune.shape_key(N, M, 1, V, 5)
    autotune.clear()
    ref = torch.zeros((V, N), dtype=torch.float64, device=dev)
    engine.phase1_xz(g, 0, M, 1000, annot, inv, 0, V, 77, ref)
    path = os.path.join(tempfile.mkdtemp(), "sched.json")
    entry = autotune.tune_phase1(g, V, reps=1, split_candidates=(1, 2, 7), range_kib_candidates=(64, 4096),
                                 path=path)
    assert entry["choice"]["p1_splits"] in (1, 2, 7)
    assert entry["choice"]["p1_finalize_kib"] in (64, 4096)
    assert entry["choice"]["p1_miss"] in (1, 2)
    assert autotune.get_value("p1_splits", key) == entry["choice"]["p1_splits"]
    # splits and finalize ranges keep the results of one missing-call method to 1e-13
    # (the two methods quantize the missing-call term separately: ~1e-12 apart)
    autotune.set_value("p1_miss", key, 1)
    autotune.set_value("p1_splits", key, 0)
    autotune.set_value("p1_finalize_kib", key, 0)
    ref.zero_()
    engine.phase1_xz(g, 0, M, 1000, annot, inv, 0, V, 77, ref)
    for s in (1, 2, 7):
        for kib in (64, 4096):
            autotune.set_value("p1_splits", key, s)
            autotune.set_value("p1_finalize_kib", key, kib)
            got = torch.zeros((V, N), dtype=torch.float64, device=dev)
            engine.phase1_xz(g, 0, M, 1000, annot, inv, 0, V, 77, got)
            scale = float(ref.abs().max())
            assert float((got - ref).abs().max()) <= 1e-13 * scale, (s, kib)
    autotune.clear()
    loaded = autotune.apply_cached(N, M, 1, V, 5, path=path)
    assert loaded is not None and loaded["choice"] == entry["choice"]
    assert autotune.get_value("p1_splits", key) == entry["choice"]["p1_splits"]
    autotune.clear()
    g.close()


@pytest.mark.parametrize("n,impute", [(1000, "mean"), (1100, "mean"), (1100, "hwe")])
def test_two_chunks_per_barrier_is_bit_identical(n, impute):
    """The two-chunks-per-barrier contraction (a schedule knob) gives bit-identical
    phase-1 and phase-2 results: even and odd phase-2 chunk counts (N = 1000 -> 8
    chunks, 1100 -> 9) and phase-1 K splits of the 23 SNP chunks of M = 2900 into
    3 (8/8/7 chunks) and 5 (5/5/5/5/3) -- splits whose last pair of chunks is a
    single chunk -- with the missing plane or not."""
    from summit_amd import autotune, engine
    dev = engine.default_device()
    M, V = 2900, 40
    print(f"cpb parity N={n} M={M} {impute}", flush=True)
    g = engine.Genotype.synthetic(n, M, seed=21, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.5)
    g.prepare(ddof=1, impute=impute, impute_seed=4)
    annot = torch.ones((M, 1), dtype=torch.float64, device=dev)
    inv = engine.resvar_inv(g, 0, M)
    k1 = autotune.shape_key(n, M, 1, V, 5)
    k2 = (M, n, k1[2], 5)
    autotune.clear()
    outs = {}
    for cpb in (1, 2):
        for splits in (1, 3, 5):
            autotune.set_value("p1_cpb", k1, cpb)
            autotune.set_value("p2_cpb", k2, cpb)
            autotune.set_value("p1_splits", k1, splits)
            xz = torch.zeros((V, n), dtype=torch.float64, device=dev)
            engine.phase1_xz(g, 0, M, 1000, annot, inv, 0, V, 99, xz)
            meansq = torch.zeros((M, 1), dtype=torch.float64, device=dev)
            engine.phase2_accum(g, 0, M, inv, V, xz, meansq, n)
            outs[(cpb, splits)] = (xz.cpu().numpy(), meansq.cpu().numpy())
    autotune.clear()
    g.close()
    ref = outs[(1, 1)]
    for key, (xz, ms) in outs.items():
        assert np.array_equal(xz, ref[0]), key
        assert np.array_equal(ms, ref[1]), key
