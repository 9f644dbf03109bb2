"""Probe tiles of the estimator replayed from one HIP graph equal eager tiles
bit for bit, replay after replay, with eager work between the replays."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.mark.parametrize("impute,probe_tile", [("mean", None), ("hwe", 16)])
def test_graphed_tiles_equal_eager_tiles(impute, probe_tile):
    from summit_amd import engine, ldscore
    dev = engine.default_device()
    N, M, T, seed = 3000, 5000, 48, 1101
    print(f"graphed tiles N={N} M={M} tile={T} {impute} probe_tile={probe_tile}", flush=True)
    g = engine.Genotype.synthetic(N, M, seed=5, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.8)
    g.prepare(ddof=1, impute=impute, impute_seed=7)
    annot = torch.ones((M, 1), dtype=torch.float64, device=dev)
    annot[::3, 0] = 0.5
    inv = engine.resvar_inv(g, 0, M)
    # eager reference after 1..4 tiles
    ref_ms = torch.zeros((M, 1), dtype=torch.float64, device=dev)
    refs = []
    for t in range(4):
        xz = torch.zeros((T, N), dtype=torch.float64, device=dev)
        engine.phase1_xz(g, 0, M, 1000, annot, inv, t * T, T, seed, xz, probe_tile=probe_tile)
        engine.phase2_accum(g, 0, M, inv, T, xz, ref_ms, N)
        refs.append(ref_ms.cpu().numpy().copy())
    ms = torch.zeros((M, 1), dtype=torch.float64, device=dev)
    gt = ldscore.GraphedTiles(g, T, seed, annot, inv, ms, probe_tile=probe_tile)
    gt.eager()
    assert np.array_equal(ms.cpu().numpy(), refs[0])
    gt.capture()
    for t in range(1, 4):
        # the program's own eager work between replays
        scratch = torch.randn(1000, device=dev).sum()
        torch.cuda.synchronize()
        gt.replay()
        torch.cuda.synchronize()
        assert np.array_equal(ms.cpu().numpy(), refs[t]), t
        assert float(scratch) == float(scratch)
    assert int(gt.v_start.item()) == 4 * T
    g.close()
