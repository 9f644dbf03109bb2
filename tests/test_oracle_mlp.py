"""Pins of the MLP oracle (CPU): BF16 rounding, GELU, the analytic gradient
against finite differences, and Adam's closed-form steps."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import mlp as om


def test_bf16_rounding_matches_torch_and_hand_ties():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-6, 6, 4000),
                        [0.0, -0.0, 1.0, -2.5, 3.0e38, 1e-40]]).astype(np.float32)
    ours = om.bf16(x)
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(ours, ref)
    # ties: 1 + 2^-8 lies halfway between 1 and 1 + 2^-7 -> even mantissa (1.0);
    # 1 + 3*2^-8 lies halfway between 1 + 2^-7 and 1 + 2^-6 -> 1 + 2^-6
    assert om.bf16(np.float32(1.0 + 2.0 ** -8)) == 1.0
    assert om.bf16(np.float32(1.0 + 3 * 2.0 ** -8)) == 1.0 + 2.0 ** -6
    assert om.bf16(np.float32(1.0 + 2.0 ** -8 + 2.0 ** -20)) == 1.0 + 2.0 ** -7


def test_gelu_known_values_and_derivative():
    x = np.array([-3.0, -1.0, 0.0, 0.5, 1.0, 2.0])
    want = np.array([0.5 * v * (1.0 + math.erf(v / math.sqrt(2.0))) for v in x])
    assert np.allclose(om.gelu(x), want, rtol=0, atol=1e-15)
    assert abs(om.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert abs(om.gelu(np.array([-1.0]))[0] + 0.15865525393145707) < 1e-15
    h = 1e-6
    num = (om.gelu(x + h) - om.gelu(x - h)) / (2 * h)
    assert np.allclose(om.gelu_grad(x), num, rtol=1e-8, atol=1e-9)


def test_analytic_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    D, R = 5, 12
    p = om.init_params(D, seed=2).astype(np.float64)
    v = om.split(p)
    v["b1"][:] = rng.standard_normal(om.HIDDEN) * 0.1
    v["b2"][:] = rng.standard_normal(om.HIDDEN) * 0.1
    v["b3"][:] = 0.3
    x = rng.standard_normal((R, D))
    t = rng.standard_normal(R)
    loss, g = om.loss_and_grad(p, x, t, mirror=False)
    # every bias and head weight, and a sample of the weight matrices
    idx = np.concatenate([
        np.arange(om.split(np.arange(om.n_params()))["b1"][0], om.split(np.arange(om.n_params()))["b1"][-1] + 1),
        rng.choice(om.HIDDEN * om.IN_PAD, 60, replace=False),
        om.HIDDEN * om.IN_PAD + om.HIDDEN + rng.choice(om.HIDDEN * om.HIDDEN, 120, replace=False),
        np.arange(om.n_params() - 2 * om.HIDDEN - 1, om.n_params()),
    ])
    h = 1e-6
    for i in idx:
        if i < om.HIDDEN * om.IN_PAD and (i % om.IN_PAD) >= D:
            continue
        pp = p.copy()
        pp[i] += h
        lp = om.loss_and_grad(pp, x, t, mirror=False)[0]
        pp[i] -= 2 * h
        lm = om.loss_and_grad(pp, x, t, mirror=False)[0]
        num = (lp - lm) / (2 * h)
        assert abs(num - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, num, g[i])
    # columns of W1 beyond the input width see zero inputs: zero gradient
    W1g = om.split(g)["W1"]
    assert np.all(W1g[:, D:] == 0.0)


def test_mirror_gradient_close_to_exact():
    rng = np.random.default_rng(3)
    D, R = 7, 200
    p = om.init_params(D, seed=4)
    x = rng.standard_normal((R, D))
    t = rng.standard_normal(R)
    l0, g0 = om.loss_and_grad(p, x, t, mirror=False)
    l1, g1 = om.loss_and_grad(p, x, t, mirror=True)
    assert abs(l1 - l0) < 2e-2 * abs(l0)
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)


def test_adam_closed_forms():
    rng = np.random.default_rng(5)
    p0 = rng.standard_normal(50)
    g = rng.standard_normal(50)
    lr, eps = 1e-2, 1e-8
    p1, m1, v1 = om.adam(p0, g, np.zeros(50), np.zeros(50), 1, lr, eps=eps)
    assert np.allclose(p1, p0 - lr * g / (np.abs(g) + eps), rtol=0, atol=1e-15)
    # second step, scalar by hand: m2 = b1 m1 + (1-b1) g2, v2 = b2 v1 + (1-b2) g2^2
    g2 = rng.standard_normal(50)
    p2, _, _ = om.adam(p1, g2, m1, v1, 2, lr, eps=eps)
    b1, b2 = 0.9, 0.999
    m2 = b1 * (1 - b1) * g + (1 - b1) * g2
    v2 = b2 * (1 - b2) * g * g + (1 - b2) * g2 * g2
    want = p1 - lr * (m2 / (1 - b1 ** 2)) / (np.sqrt(v2 / (1 - b2 ** 2)) + eps)
    assert np.allclose(p2, want, rtol=0, atol=1e-15)
    # decoupled weight decay
    p1w, _, _ = om.adam(p0, g, np.zeros(50), np.zeros(50), 1, lr, eps=eps, weight_decay=0.1)
    assert np.allclose(p1w, p0 - lr * (g / (np.abs(g) + eps) + 0.1 * p0), rtol=0, atol=1e-15)


def test_combine_weights_and_fallback():
    from summit_amd import cv
    pred = np.array([5.0, 5.0, 5.0, 5.0])
    plain = np.array([5.0, 7.0, 3.0, 500.0])
    est, lam, nfall = cv.combine(pred, plain, tau2=0.04, nvecs=8, n_snps=1000, n_samples=1001)
    s = 5.0 + 1000.0 / 1000.0
    v = 2.0 * (s * s - 1.0) / 8.0
    w = 0.04 / (0.04 + v)
    assert np.allclose(lam[:3], w)
    assert np.allclose(est[:3], pred[:3] + w * (plain[:3] - pred[:3]))
    assert nfall == 1 and est[3] == 500.0 and lam[3] == 1.0
    # many probes: the plain estimate dominates
    est2, lam2, _ = cv.combine(pred, plain, tau2=0.04, nvecs=10 ** 9, n_snps=1000, n_samples=1001)
    assert np.all(lam2[:3] > 0.999)
