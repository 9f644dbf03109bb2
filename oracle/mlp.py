"""ORACLE — CPU restatement of the learned LD-score predictor (MLP + Adam).

Test infrastructure only: ``tests/`` and ``bench.py``'s baseline leg may import
it, the product (``summit_amd``) never does.

The reference has no learned component; this network is the control variate
of the GPU engine's stochastic LD scores.  What is restated here is the
textbook algorithm the HIP kernels implement, in float64, with the GPU's
rounding points reproduced when ``mirror=True``:

* network: three linear layers ``in_dim -> 128 -> 128 -> 1``, GELU (erf form,
  ``0.5 x (1 + erf(x / sqrt 2))``) after the first two; weight matrices are
  held as FP32 masters and used as BF16 (round to nearest even), biases stay
  FP32; the inputs and both hidden activations enter the next product as
  BF16, every product accumulates in FP32 (here float64);
* loss: ``mean_r (y_r - t_r)^2``;
* backward: the chain rule of the above; the error signals that feed matrix
  products (``dz2``, ``dz1``) are BF16, bias gradients use them unrounded;
* Adam (Kingma & Ba, 2015, Algorithm 1) with bias correction and optional
  decoupled weight decay (Loshchilov & Hutter, AdamW).

Flat parameter order (FP32): ``W1 (128 x 32, row-major; inputs padded to 32
columns), b1 (128), W2 (128 x 128), b2 (128), w3 (128), b3 (1)``.

Pins (tests/test_oracle_mlp.py): the analytic gradient (``mirror=False``)
against central finite differences of the float64 forward; GELU against
``math.erf`` and its known values; BF16 rounding against torch's own
float32 -> bfloat16 conversion and hand-worked ties; Adam's first step, whose
closed form is ``p - lr * g / (|g| + eps)`` per element, and a
two-step hand computation.
"""
from __future__ import annotations

import math

import numpy as np

HIDDEN = 128
IN_PAD = 32

_SQRT1_2 = 1.0 / math.sqrt(2.0)
_INV_SQRT_2PI = 1.0 / math.sqrt(2.0 * math.pi)


def n_params(in_pad: int = IN_PAD, hidden: int = HIDDEN) -> int:
    return hidden * in_pad + hidden + hidden * hidden + hidden + hidden + 1


def split(p: np.ndarray, in_pad: int = IN_PAD, hidden: int = HIDDEN) -> dict:
    """Views of the flat parameter vector by piece."""
    p = np.asarray(p)
    o = 0
    out = {}
    for name, shape in (("W1", (hidden, in_pad)), ("b1", (hidden,)), ("W2", (hidden, hidden)), ("b2", (hidden,)),
                        ("w3", (hidden,)), ("b3", (1,))):
        size = int(np.prod(shape))
        out[name] = p[o:o + size].reshape(shape)
        o += size
    return out


def to_f32(x) -> np.ndarray:
    return np.asarray(x, dtype=np.float64).astype(np.float32)


def bf16(x) -> np.ndarray:
    """float -> float32 -> bfloat16 (round to nearest, ties to even) -> float64."""
    f = np.ascontiguousarray(to_f32(x))
    u = f.view(np.uint32).astype(np.uint64)
    nan = np.isnan(f)
    lsb = (u >> np.uint64(16)) & np.uint64(1)
    r = ((u + np.uint64(0x7FFF) + lsb) >> np.uint64(16)) << np.uint64(16)
    r = np.where(nan, np.uint64(0x7FC00000), r)
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(x: np.ndarray) -> np.ndarray:
    from scipy.special import erf
    return 0.5 * x * (1.0 + erf(x * _SQRT1_2))


def gelu_grad(x: np.ndarray) -> np.ndarray:
    from scipy.special import erf
    return 0.5 * (1.0 + erf(x * _SQRT1_2)) + x * _INV_SQRT_2PI * np.exp(-0.5 * x * x)


def init_params(in_dim: int, seed: int, in_pad: int = IN_PAD, hidden: int = HIDDEN,
                zero_head: bool = False) -> np.ndarray:
    """He-style normal init of the weight matrices (std sqrt(2 / fan_in)),
    zero biases; columns of W1 beyond in_dim are zero."""
    rng = np.random.default_rng(seed)
    p = np.zeros(n_params(in_pad, hidden), dtype=np.float64)
    v = split(p, in_pad, hidden)
    v["W1"][:, :in_dim] = rng.standard_normal((hidden, in_dim)) * math.sqrt(2.0 / in_dim)
    v["W2"][:] = rng.standard_normal((hidden, hidden)) * math.sqrt(2.0 / hidden)
    if not zero_head:
        v["w3"][:] = rng.standard_normal(hidden) * math.sqrt(1.0 / hidden)
    return to_f32(p)


def forward(p, x, mirror: bool = True, in_pad: int = IN_PAD, hidden: int = HIDDEN):
    """y (R,) and the cache for backward.  x: (R, in_dim <= in_pad)."""
    x = np.asarray(x, dtype=np.float64)
    xp = np.zeros((x.shape[0], in_pad))
    xp[:, :x.shape[1]] = x
    v = split(np.asarray(p, dtype=np.float64), in_pad, hidden)
    rw = bf16 if mirror else (lambda a: np.asarray(a, dtype=np.float64))
    W1, W2, w3 = rw(v["W1"]), rw(v["W2"]), rw(v["w3"])
    xa = rw(xp)
    z1 = xa @ W1.T + v["b1"]
    if mirror:
        z1 = to_f32(z1).astype(np.float64)
    h1 = rw(gelu(z1))
    z2 = h1 @ W2.T + v["b2"]
    if mirror:
        z2 = to_f32(z2).astype(np.float64)
    h2 = rw(gelu(z2))
    y = h2 @ w3 + v["b3"][0]
    cache = {"xa": xa, "z1": z1, "h1": h1, "z2": z2, "h2": h2, "W2": W2, "w3": w3, "mirror": mirror}
    return y, cache


def loss_and_grad(p, x, t, mirror: bool = True, in_pad: int = IN_PAD, hidden: int = HIDDEN):
    """(loss, flat gradient) of mean_r (y_r - t_r)^2."""
    y, c = forward(p, x, mirror, in_pad, hidden)
    t = np.asarray(t, dtype=np.float64)
    R = y.shape[0]
    r = y - t
    loss = float(np.sum(r * r) / R)
    dy = 2.0 * r / R
    rw = bf16 if mirror else (lambda a: a)
    g = np.zeros(n_params(in_pad, hidden))
    gv = split(g, in_pad, hidden)
    gv["b3"][0] = dy.sum()
    gv["w3"][:] = dy @ c["h2"]
    dz2 = dy[:, None] * c["w3"][None, :] * gelu_grad(c["z2"])
    gv["b2"][:] = dz2.sum(axis=0)
    dz2a = rw(dz2)
    gv["W2"][:] = dz2a.T @ c["h1"]
    dh1 = dz2a @ c["W2"]
    dz1 = dh1 * gelu_grad(c["z1"])
    gv["b1"][:] = dz1.sum(axis=0)
    dz1a = rw(dz1)
    gv["W1"][:] = dz1a.T @ c["xa"]
    return loss, g


def adam(p, g, m, v, step: int, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8,
         weight_decay: float = 0.0):
    """One Adam(W) step on float64 copies; ``step`` is 1 for the first update.
    Returns (p, m, v)."""
    p = np.asarray(p, dtype=np.float64)
    g = np.asarray(g, dtype=np.float64)
    m = beta1 * np.asarray(m, dtype=np.float64) + (1.0 - beta1) * g
    v = beta2 * np.asarray(v, dtype=np.float64) + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** step)
    vh = v / (1.0 - beta2 ** step)
    p = p - lr * (mh / (np.sqrt(vh) + eps) + weight_decay * p)
    return p, m, v
