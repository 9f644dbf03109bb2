"""Real-valued probe streams, which the library draws on the host (gpe_probes
does no device work for them), equal the oracle's MT19937-64 + polar-normal
restatement bit for bit, for both distributions and precisions, on full
1,000-SNP blocks and ragged ones.  Rademacher streams are drawn on the device
and are compared in the GPU tests."""
from __future__ import annotations

import numpy as np
import pytest


@pytest.mark.parametrize("dist", ["gaussian", "spherical"])
@pytest.mark.parametrize("fp32", [False, True])
def test_host_probe_streams_equal_oracle(dist, fp32):
    from oracle import oracle
    from summit_amd import engine
    starts = [0, 1000, 1777, 1778, 3000]
    end = 3999
    z = engine.probes(1303, starts, end, 64, 37, rand_dist=dist, fp32=fp32)
    ref = np.concatenate([oracle.probe_block(1303, s, 64, e - s, 37, dist, fp32=fp32)
                          for s, e in zip(starts, starts[1:] + [end])])
    np.testing.assert_array_equal(z, ref)


def test_host_gaussian_block_many_threads():
    """Enough blocks that the fill is spread over host threads."""
    from oracle import oracle
    from summit_amd import engine
    starts = list(range(0, 64_000, 1000))
    z = engine.probes(7, starts, 64_000, 0, 16, rand_dist="gaussian", fp32=True)
    for b in (0, 31, 63):
        ref = oracle.probe_block(7, starts[b], 0, 1000, 16, "gaussian", fp32=True)
        np.testing.assert_array_equal(z[b * 1000:(b + 1) * 1000], ref)
