"""Write the golden vectors of tests/golden/ from the reference's own code.

Run in a container that has the reference tree: it builds oracle/_ref (the
reference's src/native/common/genotype.cpp compiled from where it lies, plus
libstdc++'s std::mt19937_64 / std::normal_distribution with the reference's
seed derivation) and records inputs and outputs as small .npz fixtures.  The
fixtures are data: input genotype codes and the reference's outputs.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402
from tests.helpers import random_codes  # noqa: E402


def main() -> None:
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all", "ref"], check=True)
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_genotype.so"))
    ref.ref_read_block_standardized_double.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p]
    ref.ref_read_block_mailman_hwe.argtypes = [
        ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    ref.ref_std_probe_block.argtypes = [ctypes.c_uint64, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_void_p]

    # --- genotype decode / standardization / HWE completion
    m, n = 40, 37
    codes = random_codes(m, n, seed=2024, miss_rate=0.08, monomorphic=2, all_missing=1)
    rows = np.array([0, 2, 3, 5, 8, 13, 21, 22, 30, 34, 36], dtype=np.int64)
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, "g")
    o.write_bed(prefix, codes)
    out = {"codes": codes, "rows": rows, "blk_start": 3, "blk_end": 31, "impute_seed": 0x5EED1234ABCD}
    for mode, mi in (("mean", 0), ("hwe", 1)):
        for tag, r in (("all", np.arange(n, dtype=np.int32)), ("sub", rows.astype(np.int32))):
            X = np.empty((28, len(r)))
            rc = ref.ref_read_block_standardized_double((prefix + ".bed").encode(), (prefix + ".fam").encode(), 3, 31,
                                                        r.ctypes.data, len(r), 1, mi, out["impute_seed"], X.ctypes.data)
            assert rc == 0
            out[f"X_{mode}_{tag}"] = X.T.copy()
    r = rows.astype(np.int32)
    mean = np.empty(28)
    inv = np.empty(28)
    obs = np.empty(28, dtype=np.int32)
    rc = ref.ref_read_block_mailman_hwe((prefix + ".bed").encode(), (prefix + ".fam").encode(), 3, 31, r.ctypes.data,
                                        len(r), 1, out["impute_seed"], mean.ctypes.data, inv.ctypes.data,
                                        obs.ctypes.data)
    assert rc == 0
    out["hwe_mailman_mean_sub"] = mean
    out["hwe_mailman_inv_std_sub"] = inv
    np.savez_compressed(os.path.join(HERE, "genotype_decode.npz"), **out)

    # --- probe streams (std::mt19937_64 + std::normal_distribution<double>)
    pr = {}
    for root, blk, v0, L, V in ((1101, 0, 0, 50, 10), (1101, 1000, 64, 33, 7), (2**63 + 12345, 7, 3, 19, 5)):
        for dist, di in (("rademacher", 0), ("gaussian", 1), ("spherical", 2), ("gaussian32", 3),
                         ("spherical32", 4)):
            Z = np.empty((V, L))
            ref.ref_std_probe_block(root, blk, v0, L, V, di, Z.ctypes.data)
            pr[f"{dist}_{root}_{blk}_{v0}_{L}_{V}"] = Z.T.copy()
    np.savez_compressed(os.path.join(HERE, "probe_streams.npz"), **pr)
    print("wrote", os.listdir(HERE))


if __name__ == "__main__":
    main()
