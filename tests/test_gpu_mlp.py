"""GPU parity of the learned predictor's HIP kernels against the CPU oracle:
forward, the gradient of one step, Adam updates, training progress, and
HIP-graph replays against eager steps."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch, torch.device("cuda", 0)


def _params(D, seed):
    from oracle import mlp as om
    rng = np.random.default_rng(seed + 100)
    p = om.init_params(D, seed).astype(np.float64)
    v = om.split(p)
    v["b1"][:] = rng.standard_normal(om.HIDDEN) * 0.2
    v["b2"][:] = rng.standard_normal(om.HIDDEN) * 0.2
    v["b3"][:] = 0.1
    return p.astype(np.float32)


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("R,D", [(1000, 18), (64, 32), (37, 1)])
def test_forward_matches_oracle(torch_dev, R, D):
    torch, dev = torch_dev
    from oracle import mlp as om
    from summit_amd import mlp
    print(f"mlp forward R={R} D={D}", flush=True)
    rng = np.random.default_rng(R + D)
    x = rng.standard_normal((R, D))
    p = _params(D, 7)
    net = mlp.Net(D, max_rows=256)  # several chunks for R=1000
    net.set_params(p)
    y = net.forward(mlp.pad_features(torch.from_numpy(x).to(dev))).cpu().numpy()
    net.close()
    yo, _ = om.forward(p, x, mirror=True)
    assert _rel(y, yo) < 1e-4
    assert np.max(np.abs(y - yo)) < 2e-3 * np.max(np.abs(yo))


@pytest.mark.parametrize("R", [1000, 512, 40])
def test_gradient_matches_oracle(torch_dev, R):
    torch, dev = torch_dev
    from oracle import mlp as om
    from summit_amd import mlp
    print(f"mlp gradient R={R}", flush=True)
    D = 18
    rng = np.random.default_rng(R)
    x = rng.standard_normal((R, D))
    t = rng.standard_normal(R)
    p = _params(D, 11)
    net = mlp.Net(D, max_rows=1024)
    net.set_params(p)
    xb = mlp.pad_features(torch.from_numpy(x).to(dev))
    xbT = torch.zeros((32, (R + 63) // 64 * 64), dtype=torch.int16, device=dev)
    xbT[:, :R] = xb.t()
    tt = torch.from_numpy(t).to(dev).float()
    loss = torch.zeros(1, dtype=torch.float32, device=dev)
    net.train_step(xb, xbT, tt, mlp.Adam(lr=1e-3), update=False, loss=loss)
    g = net.get_grad()
    assert net.step_count() == 0
    assert np.array_equal(net.get_params(), p)  # no update
    lo, go = om.loss_and_grad(p, x, t, mirror=True)
    assert abs(float(loss.item()) - lo) < 1e-5 * lo
    gs, gos = om.split(g), om.split(go)
    for name in ("W1", "b1", "W2", "b2", "w3", "b3"):
        assert _rel(gs[name], gos[name]) < 2e-3, name
    net.close()


def test_adam_steps_match_oracle(torch_dev):
    torch, dev = torch_dev
    from oracle import mlp as om
    from summit_amd import mlp
    D, R = 18, 300
    rng = np.random.default_rng(3)
    x = rng.standard_normal((R, D))
    t = rng.standard_normal(R)
    p = _params(D, 12)
    net = mlp.Net(D, max_rows=512)
    net.set_params(p)
    xb = mlp.pad_features(torch.from_numpy(x).to(dev))
    xbT = torch.zeros((32, 320), dtype=torch.int16, device=dev)
    xbT[:, :R] = xb.t()
    tt = torch.from_numpy(t).to(dev).float()
    opt = mlp.Adam(lr=1e-2, weight_decay=0.01)
    m = np.zeros(p.size)
    v = np.zeros(p.size)
    pref = p.astype(np.float64)
    for step in (1, 2, 3):
        net.train_step(xb, xbT, tt, opt, update=True)
        g = net.get_grad().astype(np.float64)  # the gradient the step used
        pref, m, v = om.adam(pref, g, m, v, step, opt.lr, opt.beta1, opt.beta2, opt.eps, opt.weight_decay)
        got = net.get_params().astype(np.float64)
        assert np.max(np.abs(got - pref)) < 1e-5 * max(1.0, np.max(np.abs(pref)))
        pref = got  # continue from the device state (FP32 masters)
    assert net.step_count() == 3
    net.close()


def test_training_fits_and_graph_replays_equal_eager(torch_dev):
    torch, dev = torch_dev
    from summit_amd import mlp
    D, R = 6, 3000
    rng = np.random.default_rng(9)
    x = rng.standard_normal((R, D))
    t = np.sin(x[:, 0]) + 0.5 * x[:, 1] ** 2 - 0.3 * x[:, 2] * x[:, 3]
    t = (t - t.mean()) / t.std()
    xs = torch.from_numpy(x).to(dev).float()
    ts = torch.from_numpy(t).to(dev).float()
    p0 = mlp.init_params(D, 21, zero_head=True)
    results = []
    for graph in (True, False):
        net = mlp.Net(D, max_rows=512)
        net.set_params(p0)
        log = mlp.train(net, xs, ts, epochs=6, batch=504, opt=mlp.Adam(lr=3e-3), seed=4, graph=graph)
        results.append((net.get_params(), log.train_loss))
        net.close()
    (pg, lg), (pe, le) = results
    # every replay (one per epoch, with eager shuffles between) equals the eager epoch
    assert lg == le
    assert np.array_equal(pg, pe)
    net = mlp.Net(D, max_rows=512)
    net.set_params(p0)
    log = mlp.train(net, xs, ts, epochs=60, batch=504, opt=mlp.Adam(lr=3e-3), seed=4, graph=True)
    assert log.train_loss[-1] < 0.05 * log.train_loss[0]
    net.close()
