"""NumPy's Generator(Philox).integers(0, 2, int8) reads the high bit of each byte of
the uint32 stream (low half of each 64-bit output first, bytes little-endian) --
the rule the GPU Philox kernel implements (src/native/gxeldcore.cpp:10111-10119)."""
from __future__ import annotations

import numpy as np
import pytest


@pytest.mark.parametrize("seed", [1, 2**63 + 5, 123456789])
def test_philox_int8_rule(seed):
    L = 77
    ref = np.random.Generator(np.random.Philox(seed)).integers(0, 2, size=L, dtype=np.int8)
    raw = np.random.Philox(seed).random_raw(size=(L + 7) // 8)
    u32 = np.stack([raw & np.uint64(0xFFFFFFFF), raw >> np.uint64(32)], axis=1).reshape(-1).astype(np.uint32)
    by = u32.view(np.uint8)[:L]
    np.testing.assert_array_equal((by >> 7).astype(np.int8), ref)


def test_gxe_seed_matches_genome_wide_seed(oracle):
    from summit_amd import gxe
    for args in ((1, 0, 0), (1303, 1000, 64), (2**64 - 1, 7, 3)):
        assert gxe.make_seed(*args) == oracle.make_seed(*args)


def test_philox_keys_are_numpys():
    from summit_amd import gxe
    k = gxe.philox_keys(99, [0, 10], [3, 4])
    assert k.shape == (8,)
    np.testing.assert_array_equal(k[:2], np.random.Philox(gxe.make_seed(99, 0, 3)).state["state"]["key"])


def test_vectorized_philox_keys_match_numpy():
    """The vectorized SeedSequence key derivation against NumPy's own Philox seeding,
    over make_seed(root, block, probe) seeds and edge-case seeds (0, 2^32-1, 2^32, 2^64-1)."""
    import numpy as np
    from summit_amd import gxe
    starts, pids = [0, 1000, 7000, 123456], [0, 1, 5, 63, 1000]
    got = gxe.philox_keys(1303, starts, pids).reshape(len(starts), len(pids), 2)
    for b, s in enumerate(starts):
        for v, p in enumerate(pids):
            ref = np.random.Philox(gxe.make_seed(1303, s, p)).state["state"]["key"]
            np.testing.assert_array_equal(got[b, v], ref)
    edge = np.array([0, 1, 2**32 - 1, 2**32, 2**64 - 1, 12345678901234567], dtype=np.uint64)
    ref = np.array([np.random.Philox(int(x)).state["state"]["key"] for x in edge])
    np.testing.assert_array_equal(gxe.seed_sequence_philox_keys(edge), ref)
