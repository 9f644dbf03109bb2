"""Learned prediction of the stochastic genome-wide LD scores, combined with a
cheap plain run (a learned shrinkage estimator: the "control variate" of the
plan).

A small network predicts each SNP's LD score from features that are cheap
and exact on the GPU:

* the exact in-window LD sums over the +-50 neighbouring SNPs (unbiased r^2,
  total and per annotation), the total over +-25, and the sums over each
  side of the SNP (before / after it) at +-50 and +-25, from the banded
  int8-MFMA Gram of the windowed engine;
* the minor allele frequency;
* the annotation bits.

The network (``summit_amd.mlp``, HIP forward / backward / Adam) learns the
remainder ``(L_j - W_j) / max(W_j, 1)`` between the LD score and its exact
in-window part ``W_j`` from a high-probe-count run of the plain estimator on a training
genotype set: minimizing squared error against unbiased noisy labels fits
their conditional mean, so the prediction can be far more accurate than any
single label.

The prediction ``f_j`` is then combined with a cheap plain run with B probes
(estimate ``Lp_j``, Monte Carlo variance ``v_j``):

    L_j^cv = f_j + lam_j (Lp_j - f_j),   lam_j = tau^2 / (tau^2 + v_j),

the mean-squared-error-optimal shrinkage of an unbiased estimate towards a
prediction of squared error tau^2 (calibrated on held-out training SNPs after
removing the labels' own Monte Carlo variance).  It is biased per SNP by
(1 - lam_j)(f_j - L_j): unlike a classical control variate (which subtracts
a zero-mean correction) it trades that bias for variance, and
``CVResult.bias_bound`` reports its size, (1 - lam_j) tau averaged over SNPs
relative to the mean LD score.  As B grows, v_j -> 0, lam_j -> 1 and the
estimate becomes the plain one; a SNP whose plain estimate departs from the
prediction by more than ``guard`` standard deviations (default 10: with few
probes the plain estimate has chi-square-like tails) keeps the plain
estimate (the fallback).  ``v_j`` is the Rademacher variance
``2[(sum_i r_ij^2)^2 - sum_i r_ij^4] / B`` with ``sum_i r_ij^2`` from the
prediction (``f_j + M/(N-1)``) and the self term ``r_jj^4 = 1``
(tests/test_pgen_gwld.py:763-766 of the reference states the variance).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import engine, ldscore, mlp, winldcore

FEATURE_WINDOW = 50


def snp_features(geno: engine.Genotype, annot_bits: np.ndarray | None = None,
                 window: int = FEATURE_WINDOW) -> np.ndarray:
    """(M, 7 + 2K) FP64: in-window unbiased r^2 sums (total, then per
    annotation), the total over +-window/2, the one-sided sums (before, after)
    over window and window/2, MAF, annotation bits."""
    M = geno.m
    K = 0 if annot_bits is None else int(np.asarray(annot_bits).shape[1])
    ann = np.ones((M, 1 + K), dtype=np.float64)
    if K:
        ann[:, 1:] = np.asarray(annot_bits, dtype=np.float64)
    win = winldcore.windowed_ld(geno, 0, M, annot=ann, window_snps=int(window))
    # the total over the inner half of the window: with the full-window total it
    # describes how fast LD decays, which is what sets the part beyond the window
    inner = winldcore.windowed_ld(geno, 0, M, window_snps=max(1, int(window) // 2))
    # each side's sums: LD that decays differently before and after a SNP (a
    # haplotype-block boundary on one side) sets a different part beyond the window
    sided = [winldcore.windowed_ld(geno, 0, M, window_snps=w, side=side)
             for w in (int(window), max(1, int(window) // 2)) for side in (1, 2)]
    st = geno.stats()
    p = 0.5 * st["mean"]
    maf = np.minimum(p, 1.0 - p)
    parts = [win, inner, *sided, maf[:, None]]
    if K:
        parts.append(ann[:, 1:])
    return np.ascontiguousarray(np.concatenate(parts, axis=1))


def exact_ldscore(geno: engine.Genotype, annot: np.ndarray | None = None, ddof: int = 1) -> np.ndarray:
    """Exact (partitioned) LD scores L_jk = sum_i r_ij^2 a_ik - M_k/(N - ddof),
    M_k = sum_i a_ik, over every SNP pair: the banded int8-MFMA Gram with a
    window spanning all SNPs gives sum_i r~_ij^2 a_ik with
    r~^2 = r^2 - (1 - r^2)/(N - 2), from which
    sum_i r_ij^2 a_ik = (sum_i r~_ij^2 a_ik + M_k/(N - 2)) (N - 2)/(N - 1).
    r is the correlation of the mean-imputed genotypes, the one the stochastic
    estimator targets (its residual scales normalize each column).  Returns
    (M,) without annotations, else (M, K)."""
    M, N = geno.m, geno.n
    a = np.ones((M, 1)) if annot is None else np.asarray(annot, dtype=np.float64)
    mk = a.sum(axis=0)
    w = winldcore.windowed_ld(geno, 0, M, annot=a, window_snps=max(1, M))
    s = (w + mk[None, :] / (N - 2.0)) * (N - 2.0) / (N - 1.0)
    out = s - mk[None, :] / float(N - ddof)
    return out[:, 0] if annot is None else out


def remainder_scale(w: np.ndarray) -> np.ndarray:
    """The network learns the remainder L - W relative to max(W, 1): the
    labels' Monte Carlo noise grows with the LD score, and the error that
    matters is relative, so scaled targets have comparable noise everywhere."""
    return np.maximum(np.asarray(w, dtype=np.float64), 1.0)


@dataclass
class LDScoreModel:
    params: np.ndarray
    in_dim: int
    feat_mean: np.ndarray
    feat_sd: np.ndarray
    resid_mean: float
    resid_sd: float
    tau2: float
    tau2_raw: float
    window: int = FEATURE_WINDOW
    info: dict = field(default_factory=dict)

    def predict(self, feats: np.ndarray, net: mlp.Net | None = None) -> np.ndarray:
        """Predicted LD scores (M,) FP64 for features from ``snp_features``."""
        feats = np.asarray(feats, dtype=np.float64)
        if feats.ndim != 2 or feats.shape[1] != self.in_dim:
            raise RuntimeError("feature matrix does not match the model")
        dev = engine.default_device()
        own = net is None
        if own:
            net = mlp.Net(self.in_dim, max_rows=65536)
        try:
            net.set_params(self.params)
            xs = torch.from_numpy((feats - self.feat_mean) / self.feat_sd).to(dev)
            y = net.forward(mlp.pad_features(xs))
            g = y.to(torch.float64).cpu().numpy()
        finally:
            if own:
                net.close()
        w = feats[:, 0]
        return w + remainder_scale(w) * (self.resid_mean + self.resid_sd * g)


def fit(feats: np.ndarray, labels: np.ndarray, label_var: np.ndarray | None = None, *, epochs: int = 60,
        batch: int = 4096, lr: float = 1e-3, weight_decay: float = 1e-2, seed: int = 0, val_frac: float = 0.2,
        graph: bool = True, patience: int = 6, val_block: int = 1000) -> LDScoreModel:
    """Train the predictor on features (M, D) and per-SNP labels (M,) from a
    plain run (``label_var``: their Monte Carlo variances, for calibration)."""
    feats = np.asarray(feats, dtype=np.float64)
    labels = np.asarray(labels, dtype=np.float64)
    M, D = feats.shape
    if labels.shape != (M,):
        raise RuntimeError("one label per SNP")
    ok = np.isfinite(labels) & np.all(np.isfinite(feats), axis=1)
    # held-out SNPs are whole blocks of consecutive SNPs: labels of nearby SNPs
    # share probe noise through LD, so a SNP-level split would reward fitting it
    rng = np.random.default_rng(seed)
    n_blk = max(2, -(-M // int(val_block)))
    blk = np.arange(M) * n_blk // M
    held = rng.permutation(n_blk)[:max(1, int(round(val_frac * n_blk)))]
    is_val = np.isin(blk, held)
    val = np.flatnonzero(ok & is_val)
    tr = np.flatnonzero(ok & ~is_val)
    if val.size == 0 or tr.size == 0:
        raise RuntimeError("too few SNPs to split into training and held-out blocks")
    n_val = int(val.size)
    fm = feats[tr].mean(axis=0)
    fs = feats[tr].std(axis=0)
    fs = np.where(fs > 0, fs, 1.0)
    resid = (labels - feats[:, 0]) / remainder_scale(feats[:, 0])
    rm = float(resid[tr].mean())
    rs = float(resid[tr].std()) or 1.0
    dev = engine.default_device()
    x = torch.from_numpy((feats - fm) / fs).to(dev)
    t = torch.from_numpy((resid - rm) / rs).to(dev)
    tr_d = torch.from_numpy(tr).to(dev)
    val_d = torch.from_numpy(val).to(dev)
    net = mlp.Net(D, max_rows=batch)
    try:
        net.set_params(mlp.init_params(D, seed, zero_head=True))
        log = mlp.train(net, x[tr_d], t[tr_d], x[val_d], t[val_d], epochs=epochs, batch=batch,
                        opt=mlp.Adam(lr=lr, weight_decay=weight_decay), seed=seed, graph=graph, patience=patience)
        params = net.get_params()
    finally:
        net.close()
    model = LDScoreModel(params, D, fm, fs, rm, rs, 0.0, 0.0)
    pred = model.predict(feats[val])
    mse = float(np.mean((labels[val] - pred) ** 2))
    noise = float(np.mean(label_var[val])) if label_var is not None else 0.0
    floor = (1e-3 * float(np.mean(np.abs(labels[val])))) ** 2
    model.tau2_raw = mse - noise
    model.tau2 = max(model.tau2_raw, floor)
    model.info = {"train_snps": int(tr.size), "val_snps": int(n_val), "epochs": log.epochs, "steps": log.steps,
                  "best_epoch": log.best_epoch, "train_seconds": round(log.seconds, 3),
                  "val_loss_start": round(log.val_loss_start, 5),
                  "val_mse": mse, "label_noise_var": noise,
                  "train_loss": [round(v, 5) for v in log.train_loss],
                  "val_loss": [round(v, 5) for v in log.val_loss]}
    return model


@dataclass
class CVResult:
    ldscore: np.ndarray      # combined estimate (M,)
    prediction: np.ndarray   # network prediction f (M,)
    plain: np.ndarray        # plain B-probe estimate (M,)
    weight: np.ndarray       # lam per SNP
    fallback: int            # SNPs that kept the plain estimate
    nvecs: int
    seconds: dict
    bias_bound: float = 0.0  # mean (1 - lam) tau / mean |L|: the shrinkage's typical relative bias


def combine(pred: np.ndarray, plain: np.ndarray, tau2: float, nvecs: int, n_snps: int, n_samples: int,
            ddof: int = 1, guard: float = 10.0):
    """lam-weighted combination of prediction and plain estimate (module doc)."""
    s = np.maximum(pred, 0.0) + n_snps / float(n_samples - ddof)
    v = 2.0 * np.maximum(s * s - 1.0, 0.0) / float(nvecs)
    lam = tau2 / (tau2 + v)
    est = pred + lam * (plain - pred)
    far = ~np.isfinite(pred) | (np.abs(plain - pred) > guard * np.sqrt(v + tau2))
    est = np.where(far, plain, est)
    lam = np.where(far, 1.0, lam)
    return est, lam, int(far.sum())


def cv_ldscore(geno: engine.Genotype, model: LDScoreModel, annot_bits: np.ndarray | None = None, nvecs: int = 8,
               seed: int = 1505, step_size: int = 1000, feats: np.ndarray | None = None,
               net: mlp.Net | None = None, guard: float = 10.0) -> CVResult:
    """Control-variate LD scores of a prepared genotype handle with nvecs probes."""
    secs = {}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if feats is None:
        feats = snp_features(geno, annot_bits, model.window)
    t1 = time.perf_counter()
    pred = model.predict(feats, net)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    res = ldscore.genomewide_ldscore(geno, nvecs, seed, step_size=step_size)
    t3 = time.perf_counter()
    plain = res.ldscore[:, 0]
    est, lam, nfall = combine(pred, plain, model.tau2, nvecs, geno.m, geno.n, guard=guard)
    t4 = time.perf_counter()
    secs = {"features": t1 - t0, "predict": t2 - t1, "plain": t3 - t2, "combine": t4 - t3, "total": t4 - t0}
    bias = float(np.mean((1.0 - lam) * np.sqrt(model.tau2)) / max(float(np.mean(np.abs(est))), 1e-300))
    return CVResult(est, pred, plain, lam, nfall, int(nvecs), secs, bias)
