"""The h2 moment solve and delete-one jackknife (summit_amd.h2) against closed
forms and a simulated GWAS whose heritability is known (no GPU)."""
from __future__ import annotations

import numpy as np
import pytest

from summit_amd import h2


def test_one_bin_closed_form():
    """K = 1, a_j = 1: [T, n; n, n] sigma = [n ybar, n] with T = n^2 Lbar/M + n gives
    sigma_g = M (ybar - 1) / (n Lbar) and sigma_e = 1 - sigma_g."""
    rng = np.random.default_rng(1)
    M, n = 3000, 20000.0
    L = rng.uniform(5.0, 60.0, size=M)
    y = 1.0 + rng.gamma(2.0, 0.5, size=M)
    fit = h2.fit_h2(L, np.ones(M), y, n, h2.JackknifeDesign.blocks(M, 30))
    sg = M * (y.mean() - 1.0) / (n * L.mean())
    assert fit.h2 == pytest.approx(sg, rel=1e-12)
    assert fit.sigma[1] == pytest.approx(1.0 - sg, rel=1e-10, abs=1e-12)
    # a replicate is the same formula on the remaining SNPs, still scaled to the
    # full SNP count (the trace divides by the replicate's and the full mass)
    d = h2.JackknifeDesign.blocks(M, 30)
    keep = np.ones(M, bool)
    keep[d.starts[4]:d.ends[4]] = False
    assert fit.reps_h2[4] == pytest.approx(M * (y[keep].mean() - 1.0) / (n * L[keep].mean()), rel=1e-11)


def test_jackknife_se_textbook_and_block_means():
    """Equal blocks: SE^2 = (B-1)/B sum (theta_(-b) - mean)^2.  Unequal blocks with
    theta the overall mean: the pseudovalues are the block means exactly."""
    rng = np.random.default_rng(2)
    B = 17
    reps = rng.standard_normal(B + 1)
    sizes = np.full(B, 40.0)
    se = h2.jackknife_se(reps[:, None], sizes)[0]
    loo = reps[:B]
    # with equal sizes the pseudovalues give the textbook delete-one formula
    # around the pseudovalue mean, B theta - (B-1) mean(theta_(-b))
    assert se == pytest.approx(np.sqrt((B - 1) / B * np.sum((loo - loo.mean()) ** 2)), rel=1e-12)
    x = rng.standard_normal(500)
    sizes = np.array([30, 70, 120, 80, 200], dtype=float)
    edges = np.concatenate([[0], np.cumsum(sizes)]).astype(int)
    full = x.mean()
    loo = np.array([np.delete(x, np.arange(edges[u], edges[u + 1])).mean() for u in range(5)])
    pv = h2._pseudovalues(np.concatenate([loo, [full]])[:, None], sizes)[:, 0]
    bm = np.array([x[edges[u]:edges[u + 1]].mean() for u in range(5)])
    np.testing.assert_allclose(pv, bm, rtol=1e-10, atol=1e-12)
    w = sizes / sizes.sum()
    expect = np.sqrt(np.sum(w * w / (1 - w) * (bm - np.sum(w * bm)) ** 2))
    assert h2.jackknife_se(np.concatenate([loo, [full]])[:, None], sizes)[0] == pytest.approx(expect, rel=1e-10)


def _simulate(rng, N, M, h2_bins, annot, blocks=50):
    """Block-LD genotypes, phenotype with per-SNP effect variance set by the
    annotation bins, GWAS by OLS; exact null-subtracted LD scores."""
    hap = np.empty((2, N, M), dtype=np.int8)
    maf = rng.uniform(0.05, 0.5, size=M)
    for h in range(2):
        cur = rng.random(N) < maf[0]
        for j in range(M):
            if j % 20 == 0:
                cur = rng.random(N) < maf[j]
            else:
                redraw = rng.random(N) > 0.8
                cur = np.where(redraw, rng.random(N) < maf[j], cur)
            hap[h, :, j] = cur
    G = (hap[0] + hap[1]).astype(np.float64)
    X = (G - G.mean(0)) / G.std(0)
    per_snp = annot @ (np.asarray(h2_bins) / annot.sum(axis=0))
    beta = rng.standard_normal(M) * np.sqrt(per_snp)
    g = X @ beta
    e = rng.standard_normal(N) * np.sqrt(1.0 - sum(h2_bins))
    ph = g + e
    ph = ph - ph.mean()
    xx = np.sum(X * X, axis=0)
    b = X.T @ ph / xx
    resid = np.sum(ph * ph) - b * b * xx
    se = np.sqrt(resid / (N - 2) / xx)
    r = X.T @ X / N
    L = (r * r) @ annot - annot.sum(axis=0)[None, :] / (N - 1)
    return b, se, L


def test_simulated_h2_recovered_within_jackknife_se():
    rng = np.random.default_rng(3)
    N, M = 3000, 1600
    annot = np.ones((M, 1))
    b, se, L = _simulate(rng, N, M, [0.5], annot)
    y = h2.h2_response(b, se, N, N)
    fit = h2.fit_h2(L, annot, y, h2.effective_n_scale(N), h2.JackknifeDesign.blocks(M, 40))
    assert 0.01 < fit.h2_se < 0.2
    assert abs(fit.h2 - 0.5) < 4.0 * fit.h2_se


def test_partitioned_h2_and_enrichment():
    rng = np.random.default_rng(4)
    N, M = 3000, 1600
    a1 = (rng.random(M) < 0.3).astype(float)
    annot = np.column_stack([np.ones(M), a1])
    # bin 0 covers every SNP, bin 1 a 30% subset carrying half of the heritability
    b, se, L = _simulate(rng, N, M, [0.25, 0.25], annot)
    y = h2.h2_response(b, se, N, N)
    fit = h2.fit_h2(L, annot, y, h2.effective_n_scale(N), h2.JackknifeDesign.blocks(M, 40))
    assert abs(fit.h2 - 0.5) < 4.0 * fit.h2_se
    # the subset's heritability: its own 0.25 plus its share of bin 0's
    truth_1 = 0.25 + 0.25 * a1.sum() / M
    assert abs(fit.h2_cat[1] - truth_1) < 4.0 * fit.h2_cat_se[1]
    assert fit.enrichment[1] > 1.0
    assert np.isfinite(fit.enrichment_se).all()


def test_block_designs():
    d = h2.JackknifeDesign.blocks(103, 10)
    assert d.starts[0] == 0 and d.ends[-1] == 103 and np.all(d.ends > d.starts)
    L = np.ones((100, 1))
    L[:50] = 9.0
    w = h2.JackknifeDesign.blocks(100, 10, balance="ldscore", ldscores=L, min_block_snps=2)
    sizes = w.ends - w.starts
    # the first 50 SNPs hold 90% of the mass: nine blocks cover them, the last block the rest
    assert sizes[:9].sum() == 50 and sizes[-1] == 50 and sizes.min() >= 2
    c = h2.JackknifeDesign.chromosomes(np.repeat([1, 2, 5], [10, 20, 7]))
    np.testing.assert_array_equal(c.starts, [0, 10, 30])
    with pytest.raises(ValueError):
        h2.JackknifeDesign.chromosomes(np.array([2, 1]))


def test_reads_the_engine_files(tmp_path):
    pd = pytest.importorskip("pandas")
    from summit_amd import ldscore
    M = 50
    snp = pd.DataFrame({"CHR": np.repeat([1, 2], 25), "SNP": [f"rs{j}" for j in range(M)], "BP": np.arange(M) * 10})
    L = np.random.default_rng(5).uniform(1, 20, size=(M, 2))
    out = str(tmp_path / "x")
    ldscore.write_gw_ldscore_table(out, snp, pd.DataFrame(L, columns=["base", "a1"]))
    ldscore.write_gw_annotation_masses(out, [50.0, 12.0])
    chrom, snps, Lr, bins = h2.read_gw_ldscores(out)
    assert bins == ["base", "a1"] and snps[3] == "rs3" and chrom[30] == 2
    np.testing.assert_allclose(Lr, L, atol=5e-7)
    np.testing.assert_allclose(h2.read_gw_annotation_masses(out), [50.0, 12.0])
