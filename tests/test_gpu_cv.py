"""The learned control variate on the GPU: exact LD scores from the full-window
Gram against dense FP64, and what the trained network gains over the plain
estimator at the same probe count."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _dense_exact(g, torch, dev):
    """sum_i r_ij^2 - M/(N-1) from a dense FP64 product of the mean-imputed
    standardized genotypes (the estimator's residual scales)."""
    from summit_amd import engine
    M, N = g.m, g.n
    st = g.stats()
    codes = torch.from_numpy(g.download("snp_raw")).to(dev)
    sh = torch.tensor([0, 2, 4, 6], device=dev, dtype=torch.uint8)
    c = ((codes[:, :, None] >> sh) & 3).reshape(M, -1)[:, :N]
    val = torch.tensor([0.0, 0.0, 1.0, 2.0], device=dev, dtype=torch.float64)[c.long()]
    obs = (c != 1).to(torch.float64)
    x = (val - torch.from_numpy(st["mean"]).to(dev)[:, None]) * torch.from_numpy(st["inv_sd"]).to(dev)[:, None] * obs
    inv = engine.resvar_inv(g, 0, M)
    xs = x * inv[:, None]
    r = (xs @ xs.T) / float(N - 1)
    return ((r * r).sum(dim=1) - M / float(N - 1)).cpu().numpy()


def _data(seed, N, M, block_len=100, copy_prob=0.9):
    from summit_amd import engine
    g = engine.Genotype.synthetic(N, M, seed=seed, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=block_len,
                                  copy_prob=copy_prob)
    g.prepare(ddof=1, impute="mean")
    annot = (np.random.default_rng(seed + 7).random((M, 8)) < 0.1).astype(np.float64)
    return g, annot


def test_exact_ldscore_full_window_matches_dense():
    from summit_amd import cv
    dev = torch.device("cuda", 0)
    print("exact LD N=1500 M=2500", flush=True)
    g, _ = _data(11, 1500, 2500)
    got = cv.exact_ldscore(g)
    ref = _dense_exact(g, torch, dev)
    g.close()
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-9)


def test_control_variate_beats_plain_at_same_probes():
    from summit_amd import cv, ldscore
    N, M = 3000, 6000
    print(f"control variate N={N} M={M}: train on seed 505 (1000-probe labels), evaluate seed 506", flush=True)
    ga, annot_a = _data(505, N, M)
    fa = cv.snp_features(ga, annot_a)
    assert fa.shape == (M, 23)
    lab = ldscore.genomewide_ldscore(ga, 1000, seed=1505, vtiles=ldscore.bounded_vtile_sizes(1000, 512), mc=True)
    model = cv.fit(fa, lab.ldscore[:, 0], lab.mc_var[:, 0], epochs=40, batch=1024, seed=0)
    ga.close()
    gb, annot_b = _data(506, N, M)
    exact = cv.exact_ldscore(gb)
    fb = cv.snp_features(gb, annot_b)
    res = cv.cv_ldscore(gb, model, annot_b, nvecs=8, seed=1506, feats=fb)
    gb.close()

    def err(est):
        return float(np.mean(np.abs(est - exact) / np.abs(exact)))

    e_plain, e_cv, e_pred = err(res.plain), err(res.ldscore), err(res.prediction)
    e_win = err(fb[:, 0])  # the exact in-window sum alone, for the record
    print(f"mean rel err: plain B=8 {e_plain:.4f}, control variate B=8 {e_cv:.4f}, prediction {e_pred:.4f}, "
          f"in-window sum {e_win:.4f}; "
          f"tau2 {model.tau2:.3g} (raw {model.tau2_raw:.3g}), fallback {res.fallback}, "
          f"epochs {model.info['epochs']} best {model.info['best_epoch']} val start {model.info['val_loss_start']} "
          f"val {model.info['val_loss']}", flush=True)
    assert e_cv < 0.05
    assert e_cv < e_plain / 5.0
    assert res.fallback < M // 100


def test_network_learns_ld_beyond_the_window():
    """LD that reaches far past +-50 SNPs (1000-SNP haplotype blocks, copy
    probability 0.99: the part of the LD score beyond the window is about half
    of the in-window part).  Trained on labels from a 32,768-probe plain run of
    the method itself, the network must cut the mean relative error of its own
    starting point (the in-window sum plus the mean remainder) by at least 15%,
    and the combination with an 8-probe plain run must beat the plain estimator
    at the same probe count by a wide margin."""
    from summit_amd import cv, ldscore
    N, M = 3000, 20000
    print(f"long-range LD N={N} M={M}", flush=True)
    ga, annot_a = _data(515, N, M, block_len=1000, copy_prob=0.99)
    fa = cv.snp_features(ga, annot_a)
    lab = ldscore.genomewide_ldscore(ga, 32768, seed=1515, vtiles=ldscore.bounded_vtile_sizes(32768, 512), mc=True,
                                     traversal_width=4096)
    model = cv.fit(fa, lab.ldscore[:, 0], lab.mc_var[:, 0], epochs=120, batch=1024, lr=3e-4, seed=0, patience=20)
    ga.close()
    gb, annot_b = _data(516, N, M, block_len=1000, copy_prob=0.99)
    exact = cv.exact_ldscore(gb)
    fb = cv.snp_features(gb, annot_b)
    res = cv.cv_ldscore(gb, model, annot_b, nvecs=8, seed=1516, feats=fb)
    gb.close()

    def err(est):
        return float(np.mean(np.abs(est - exact) / np.abs(exact)))

    e_base = err(fb[:, 0] + cv.remainder_scale(fb[:, 0]) * model.resid_mean)
    e_pred, e_cv, e_plain = err(res.prediction), err(res.ldscore), err(res.plain)
    print(f"mean rel err: in-window + mean remainder {e_base:.4f}, network {e_pred:.4f} "
          f"(x{e_pred / e_base:.3f}), combined B=8 {e_cv:.4f}, plain B=8 {e_plain:.4f}; best epoch "
          f"{model.info['best_epoch']}, shrinkage bias bound {res.bias_bound:.4f}", flush=True)
    assert e_pred <= 0.85 * e_base
    assert e_cv < e_plain / 10.0
