"""The oracle's Mailman restatement (the reference's CPU path for at most 10
probes, timed as a CPU baseline) is pinned by the dense integer products:
exact int64 equality for G.Z and G^T U with missing calls as 0, ragged last
segments, and the segment-size rule's closed form."""
from __future__ import annotations

import math

import numpy as np
import pytest


def _codes(m, n, seed, miss):
    rng = np.random.default_rng(seed)
    maf = rng.uniform(0.05, 0.5, size=m)
    g = rng.binomial(2, maf[:, None], size=(m, n))
    codes = np.choose(g, [0, 2, 3]).astype(np.uint8)  # 0 -> 00, 1 -> 10, 2 -> 11
    codes[rng.random((m, n)) < miss] = 1              # 01 = missing
    masked = np.where(codes == 1, 0, g)
    return codes, masked


def test_segment_size_rule():
    from oracle import oracle as o
    assert o.mailman_segment_size(27) == 1
    assert o.mailman_segment_size(5000) == math.floor(math.log(5000) / math.log(3)) - 1 == 6
    assert o.mailman_segment_size(100_000) == 8   # floor(log3 1e5) - 1 = 9, capped at 8
    assert o.mailman_segment_size(81) == 3        # log3(81) = 4 exactly


@pytest.mark.parametrize("n,m,b,miss", [(30, 17, 3, 0.0), (500, 101, 10, 0.02), (2187, 64, 5, 0.1),
                                        (1, 9, 2, 0.0)])
def test_mailman_products_equal_dense_integer_products(n, m, b, miss):
    from oracle import oracle as o
    codes, masked = _codes(m, n, 7 + n, miss)
    packed = o.pack_codes(codes)
    rng = np.random.default_rng(3)
    z = (rng.integers(0, 2, size=(m, b)) * 2 - 1).astype(np.int8)
    u = o.mailman_gz(packed, n, z, threads=2)
    np.testing.assert_array_equal(u, o.gz_int(masked, z))
    w = o.mailman_gtu(packed, n, u, threads=3)
    np.testing.assert_array_equal(w, o.gtu_int(masked, u))
    # brute force on a corner
    assert u[0, 0] == sum(int(masked[j, 0]) * int(z[j, 0]) for j in range(m))
