"""Host side of the learned LD-score predictor (libsummit_gpe ``gpe_mlp_*``).

``Net`` owns one network on the GPU: FP32 master parameters, Adam moments,
BF16 weight copies and a fixed workspace for ``max_rows`` rows per training
step.  Forward, backward and Adam are the library's HIP kernels; torch only
holds the feature / target tensors and does the per-epoch shuffling.

``train`` runs minibatch Adam; with ``graph=True`` each epoch's steps are
captured once in a HIP graph (torch.cuda.CUDAGraph over the library's
launches on the capturing stream) and replayed, so an epoch costs one graph
launch instead of ten kernel launches per minibatch.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _native
from ._native import c_i32, c_i64, c_vp, call

IN_PAD = 32
HIDDEN = 128


def n_params() -> int:
    n = c_i64()
    call("gpe_mlp_n_params", ctypes.byref(n))
    return int(n.value)


def bf16_bits(x: torch.Tensor) -> torch.Tensor:
    """float tensor -> bf16 bit patterns (int16 view), round to nearest even."""
    return x.to(torch.float32).to(torch.bfloat16).view(torch.int16)


def pad_features(x: torch.Tensor) -> torch.Tensor:
    """(R, D<=32) float -> (R, 32) bf16 bits, zero beyond D."""
    if x.dim() != 2 or x.shape[1] > IN_PAD:
        raise RuntimeError("features must be (rows, <= 32)")
    out = torch.zeros((x.shape[0], IN_PAD), dtype=torch.float32, device=x.device)
    out[:, :x.shape[1]] = x
    return bf16_bits(out).contiguous()


def init_params(in_dim: int, seed: int, zero_head: bool = False) -> np.ndarray:
    """He-normal weight matrices (std sqrt(2 / fan_in)), zero biases, zero W1
    columns beyond in_dim; ``zero_head`` starts the output layer at zero so the
    untrained network predicts 0 (the mean of standardized targets)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(HIDDEN * IN_PAD + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1, dtype=np.float64)
    o = 0
    W1 = p[o:o + HIDDEN * IN_PAD].reshape(HIDDEN, IN_PAD)
    o += HIDDEN * IN_PAD + HIDDEN
    W2 = p[o:o + HIDDEN * HIDDEN].reshape(HIDDEN, HIDDEN)
    o += HIDDEN * HIDDEN + HIDDEN
    w3 = p[o:o + HIDDEN]
    W1[:, :in_dim] = rng.standard_normal((HIDDEN, in_dim)) * np.sqrt(2.0 / in_dim)
    W2[:] = rng.standard_normal((HIDDEN, HIDDEN)) * np.sqrt(2.0 / HIDDEN)
    if not zero_head:
        w3[:] = rng.standard_normal(HIDDEN) * np.sqrt(1.0 / HIDDEN)
    return p.astype(np.float32)


def _ptr(t) -> int:
    return 0 if t is None else int(t.data_ptr())


def _stream() -> int:
    return int(torch.cuda.current_stream().cuda_stream)


@dataclass
class Adam:
    lr: float = 1e-3
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    weight_decay: float = 0.0


class Net:
    def __init__(self, in_dim: int, max_rows: int = 4096):
        lib = _native.load()
        h = c_vp()
        _native.check(lib.gpe_mlp_create(int(in_dim), int(max_rows), ctypes.byref(h)))
        self._h = h
        self.in_dim = int(in_dim)
        self.max_rows = int(max_rows)
        self.n_params = n_params()

    # ------------------------------------------------------------ parameters
    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(np.asarray(p, dtype=np.float32))
        if p.size != self.n_params:
            raise RuntimeError("parameter count mismatch")
        call("gpe_mlp_set_params", self._h, int(p.ctypes.data), p.size)

    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, dtype=np.float32)
        call("gpe_mlp_get_params", self._h, int(p.ctypes.data), p.size)
        return p

    def get_grad(self) -> np.ndarray:
        g = np.empty(self.n_params, dtype=np.float32)
        call("gpe_mlp_get_grad", self._h, int(g.ctypes.data), g.size)
        return g

    def step_count(self) -> int:
        s = c_i32()
        call("gpe_mlp_step_count", self._h, ctypes.byref(s))
        return int(s.value)

    # ------------------------------------------------------------ compute
    def forward(self, xb: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        """xb: (R, 32) bf16 bits (``pad_features``); returns (R,) FP32."""
        if xb.dtype != torch.int16 or xb.dim() != 2 or xb.shape[1] != IN_PAD or not xb.is_contiguous():
            raise RuntimeError("features must be contiguous (rows, 32) bf16 bits")
        R = xb.shape[0]
        if out is None:
            out = torch.empty(R, dtype=torch.float32, device=xb.device)
        call("gpe_mlp_forward", self._h, _ptr(xb), IN_PAD, R, _ptr(out), _stream())
        return out

    def train_step(self, xb: torch.Tensor, xbT: torch.Tensor, t: torch.Tensor, opt: Adam, update: bool = True,
                   loss: torch.Tensor | None = None, rows: int | None = None, row0: int = 0) -> None:
        """One step on rows [row0, row0 + rows) of xb (R, 32), xbT (32, ldT), t (R,)."""
        R = xb.shape[0] if rows is None else int(rows)
        if xbT.shape[0] != IN_PAD or xbT.stride(1) != 1 or xbT.stride(0) % 8:
            raise RuntimeError("transposed features must be (32, ldT) with ldT a multiple of 8")
        if (row0 % 8) != 0:
            raise RuntimeError("row0 must be a multiple of 8")
        call("gpe_mlp_train_step", self._h, _ptr(xb) + row0 * IN_PAD * 2, IN_PAD, _ptr(xbT) + row0 * 2,
             xbT.stride(0), _ptr(t) + row0 * 4, R, float(opt.lr), float(opt.beta1), float(opt.beta2),
             float(opt.eps), float(opt.weight_decay), int(bool(update)), _ptr(loss), _stream())

    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            _native.load().gpe_mlp_free(self._h)
            self._h = c_vp()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class TrainLog:
    epochs: int = 0
    steps: int = 0
    seconds: float = 0.0
    train_loss: list = field(default_factory=list)
    val_loss: list = field(default_factory=list)
    val_loss_start: float = float("nan")
    best_epoch: int = -1


def train(net: Net, x: torch.Tensor, t: torch.Tensor, x_val: torch.Tensor | None = None,
          t_val: torch.Tensor | None = None, epochs: int = 40, batch: int = 4096, opt: Adam | None = None,
          seed: int = 0, graph: bool = True, patience: int = 8) -> TrainLog:
    """Minibatch Adam on standardized features x (R, D) FP32 and targets t (R,).
    Keeps the parameters with the lowest validation loss (when a validation set
    is given; the starting parameters included, best_epoch -1) and stops after
    ``patience`` epochs without a new best."""
    import time

    opt = opt or Adam()
    dev = x.device
    R = x.shape[0]
    batch = int(min(batch, net.max_rows))
    batch = max(8, batch // 8 * 8)
    xb_all = pad_features(x)
    t_all = t.to(torch.float32).contiguous()
    n_steps = (R + batch - 1) // batch
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed))
    xs = torch.empty_like(xb_all)
    ts = torch.empty_like(t_all)
    xsT = torch.zeros((IN_PAD, (R + 63) // 64 * 64), dtype=torch.int16, device=dev)
    loss_buf = torch.zeros(n_steps, dtype=torch.float32, device=dev)
    xv = None if x_val is None else pad_features(x_val)
    tv = None if t_val is None else t_val.to(torch.float32)
    log = TrainLog()

    def shuffle():
        perm = torch.randperm(R, generator=gen, device=dev)
        torch.index_select(xb_all, 0, perm, out=xs)
        torch.index_select(t_all, 0, perm, out=ts)
        xsT[:, :R].copy_(xs.t())

    def epoch_steps():
        for i in range(n_steps):
            r0 = i * batch
            rows = min(batch, R - r0)
            net.train_step(xs, xsT, ts, opt, update=True, loss=loss_buf[i:i + 1], rows=rows, row0=r0)

    def val_loss():
        yv = net.forward(xv)
        return float(torch.mean((yv - tv) ** 2).item())

    g = None
    best = (float("inf"), None)
    bad = 0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if xv is not None:
        # the starting point competes too: training that never beats it on
        # held-out rows leaves the parameters unchanged
        log.val_loss_start = val_loss()
        best = (log.val_loss_start, net.get_params())
    for ep in range(int(epochs)):
        shuffle()
        if graph:
            if g is None:
                side = torch.cuda.Stream()
                side.wait_stream(torch.cuda.current_stream())
                g = torch.cuda.CUDAGraph()
                with torch.cuda.graph(g, stream=side):
                    epoch_steps()
                torch.cuda.current_stream().wait_stream(side)
            g.replay()
        else:
            epoch_steps()
        log.train_loss.append(float(loss_buf.mean().item()))
        log.epochs = ep + 1
        log.steps += n_steps
        if xv is not None:
            vl = val_loss()
            log.val_loss.append(vl)
            if vl < best[0]:
                best = (vl, net.get_params())
                log.best_epoch = ep
                bad = 0
            else:
                bad += 1
                if bad >= patience:
                    break
    torch.cuda.synchronize()
    log.seconds = time.perf_counter() - t0
    if best[1] is not None:
        net.set_params(best[1])
    return log
