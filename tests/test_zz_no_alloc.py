"""gpe_project_f64_noalloc against gpe_project_f64 through the C ABI: the same
kernels on the same inputs give the same bits; two calls in a row (the second
smaller) show that the shared fixed scratch carries nothing between calls; sizes
beyond the fixed scratch are refused with a message."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _call(name, x, A, B, p, center):
    import torch

    from summit_amd import _native
    Q, N = x.shape
    stream = torch.cuda.current_stream().cuda_stream
    a_ptr = 0 if A is None else A.data_ptr()
    b_ptr = 0 if B is None else B.data_ptr()
    _native.call(name, x.data_ptr(), x.stride(0), N, Q, a_ptr, N, b_ptr, N, p, int(center), stream)
    torch.cuda.synchronize()
    return x


def _case(torch, dev, rng, N, Q, p, center):
    x = torch.from_numpy(rng.standard_normal((Q, N))).to(dev)
    if p == 0:
        return x, None, None
    C, _ = np.linalg.qr(rng.standard_normal((N, p)))
    A = torch.from_numpy(np.ascontiguousarray(C.T)).to(dev)      # (p, N): column k of C at k*N
    B = torch.from_numpy(np.ascontiguousarray(C.T)).to(dev)      # (p, N) row-major
    return x, A, B


def test_noalloc_projection_matches_allocating_entry():
    import torch

    from summit_amd import engine
    dev = engine.default_device()
    rng = np.random.default_rng(12)
    # a large call, then a smaller one: every output of the second depends only
    # on its own inputs although the scratch still holds the first call's values
    for (N, Q, p, center) in ((3001, 700, 20, True), (517, 33, 3, False), (257, 5, 0, True)):
        print(f"no-alloc projection: N={N} Q={Q} p={p} center={center}", flush=True)
        x, A, B = _case(torch, dev, rng, N, Q, p, center)
        ref = _call("gpe_project_f64", x.clone(), A, B, p, center).cpu().numpy()
        got = _call("gpe_project_f64_noalloc", x.clone(), A, B, p, center).cpu().numpy()
        np.testing.assert_array_equal(got, ref)
        # and both are the projection itself
        xn = x.cpu().numpy()
        if p:
            Cn = A.cpu().numpy().T
            xn = xn - (Cn @ (Cn.T @ xn.T)).T
        if center:
            xn = xn - xn.mean(axis=1, keepdims=True)
        np.testing.assert_allclose(got, xn, rtol=1e-12, atol=1e-12)


def test_noalloc_projection_refuses_sizes_beyond_its_scratch():
    import torch

    from summit_amd import _native, engine
    dev = engine.default_device()
    x = torch.zeros((2, 8), dtype=torch.float64, device=dev)
    lib = _native.load()
    # (p + 2) * n_cols + p = 66 * 8000 + 64 > 524288 doubles: refused before any launch
    rc = lib.gpe_project_f64_noalloc(ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(8), ctypes.c_int64(8),
                                     ctypes.c_int32(8000), ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(8),
                                     ctypes.c_void_p(x.data_ptr()), ctypes.c_int64(8), ctypes.c_int32(64),
                                     ctypes.c_int32(0), None)
    assert rc != 0
    assert b"scratch doubles" in lib.gpe_last_error()
