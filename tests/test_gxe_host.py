"""GxE host-side probe draws (no GPU): NumPy Philox per (SNP block, probe id)
exactly as src/summit/ldscore/gwe_ldscore.py:2484-2505 draws them."""
from __future__ import annotations

import numpy as np
import pytest

from summit_amd import gxe


@pytest.mark.parametrize("dist", ["gaussian", "spherical"])
def test_host_probes_are_numpy_philox_streams(dist):
    M, step, ids = 530, 200, [3, 4, 9]
    z = gxe.probes_host(41, 0, M, step, ids, dist)
    for s in range(0, M, step):
        e = min(M, s + step)
        for c, pid in enumerate(ids):
            v = np.random.Generator(np.random.Philox(gxe.make_seed(41, s, pid))).standard_normal(e - s)
            if dist == "spherical":
                v *= np.sqrt(e - s) / np.linalg.norm(v)
                assert np.linalg.norm(z[s:e, c]) == pytest.approx(np.sqrt(e - s), rel=1e-13)
            np.testing.assert_array_equal(z[s:e, c], v)
    z32 = gxe.probes_host(41, 0, M, step, ids, dist, fp32=True)
    np.testing.assert_array_equal(z32, z.astype(np.float32).astype(np.float64))
    with pytest.raises(ValueError, match="Unsupported rand_dist"):
        gxe.probes_host(41, 0, M, step, ids, "rademacher")
