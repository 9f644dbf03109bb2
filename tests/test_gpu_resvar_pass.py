"""The drop-in's fused residual-variance pass with the fourth-moment statistics
(compute_mu22) and the SNP-variance share (snp_var_out) against the oracle
(src/native/gwldcore.cpp:3335-3566)."""
from __future__ import annotations

import os
import tempfile

import numpy as np
import pytest

from tests.helpers import random_codes

pytestmark = pytest.mark.gpu


def _write(oracle, codes, name="rv"):
    d = tempfile.mkdtemp()
    prefix = os.path.join(d, name)
    oracle.write_bed(prefix, codes)
    return prefix


@pytest.mark.parametrize("impute,p", [("mean", 0), ("mean", 3), ("hwe", 0), ("hwe", 3), ("mean", 13)])
def test_fused_residual_pass_matches_oracle(oracle, impute, p):
    from summit_amd import gwldcore as gw
    m, n, step, seed = 3000, 257, 400, 91
    print(f"fused residual pass M={m} N={n} impute={impute} covariates={p}", flush=True)
    codes = random_codes(m, n, seed=17 + p, miss_rate=0.02, monomorphic=3, all_missing=1)
    prefix = _write(oracle, codes)
    rng = np.random.default_rng(p + 1)
    B = 10  # two passes of the kernel's bin groups (8 + 2)
    A = np.column_stack([np.ones(m), (rng.random((m, B - 2)) < 0.3).astype(float), rng.uniform(0.0, 2.0, m)])
    C = R = None
    if p:
        Q, _ = np.linalg.qr(np.column_stack([np.ones(n), rng.standard_normal((n, p - 1))]))
        C, R = np.asfortranarray(Q), np.ascontiguousarray(Q.T)
    v = np.zeros(m)
    inv, mu = gw.precompute_residual_variances_bed(prefix, prefix + ".fam", m, step, None, 1, 1e-10, True, A, C, R,
                                                   impute, 5, v)
    inv_o, mu_o, v_o = oracle.resvar_pass(oracle.pack_codes(codes), n, m, step, None, 1, 1e-10, A, C, R, impute, 5)
    np.testing.assert_allclose(inv, inv_o, rtol=1e-9 if p else 1e-12)
    np.testing.assert_allclose(mu, mu_o, rtol=1e-10)
    np.testing.assert_allclose(v, v_o, rtol=1e-9, atol=1e-12)
    # plain inverse scales without the statistics are unchanged
    inv2, none = gw.precompute_residual_variances_bed(prefix, prefix + ".fam", m, step, None, 1, 1e-10, False, None,
                                                      C, R, impute, 5)
    assert none is None
    np.testing.assert_allclose(inv2, inv, rtol=1e-13)


def test_mu22_known_answer_and_errors(oracle):
    """0, missing, 2, 2 under mean imputation: mu22 = 9/8 (x^2 = 4/3, 0, 1/3, 1/3, inv^2 = 3/2)."""
    from summit_amd import gwldcore as gw
    codes = oracle.geno_to_codes(np.array([[0, -9, 2, 2]]))
    prefix = _write(oracle, codes, "ka")
    v = np.zeros(1)
    inv, mu = gw.precompute_residual_variances_bed(prefix, prefix + ".fam", 1, 1, None, 1, 1e-10, True,
                                                   np.ones((1, 1)), None, None, "mean", None, v)
    assert inv[0] ** 2 == pytest.approx(1.5, rel=1e-13)
    assert mu[0, 0] == pytest.approx(9.0 / 8.0, rel=1e-13)
    assert v[0] == 1.0
    with pytest.raises(RuntimeError, match="annot_all must be provided"):
        gw.precompute_residual_variances_bed(prefix, prefix + ".fam", 1, 1, compute_mu22=True, impute_mode="mean")
    with pytest.raises(RuntimeError, match="snp_var_out shape mismatch"):
        gw.precompute_residual_variances_bed(prefix, prefix + ".fam", 1, 1, impute_mode="mean",
                                             snp_var_out=np.zeros(2))
