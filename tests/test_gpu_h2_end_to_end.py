"""End to end on the device: block-Markov genotypes, partitioned stochastic LD
scores from the engine written to and read back from the .gw.ldscore.gz /
.gw.M files, a simulated GWAS with known partitioned heritability, and the
moment solve with its block jackknife recovering it."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_h2_recovered_from_engine_ldscores(tmp_path):
    print("h2 end to end: N=6000 M=12000 K=3", flush=True)
    import pandas as pd

    from summit_amd import engine, h2, ldscore, simulate
    N, M = 6000, 12000
    g = engine.Genotype.synthetic(N, M, seed=202, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    rng = np.random.default_rng(7)
    annot = np.column_stack([np.ones(M), (rng.random(M) < 0.1).astype(float), (rng.random(M) < 0.1).astype(float)])
    res = ldscore.genomewide_ldscore(g, 200, seed=1202, annot=annot)
    snp = pd.DataFrame({"CHR": np.ones(M, dtype=int), "SNP": [f"rs{j}" for j in range(M)], "BP": np.arange(M) * 100})
    out = str(tmp_path / "c2")
    ldscore.write_gw_ldscore_table(out, snp, pd.DataFrame(res.ldscore, columns=["base", "a1", "a2"]))
    ldscore.write_gw_annotation_masses(out, annot.sum(axis=0))
    _, _, L, bins = h2.read_gw_ldscores(out)
    assert bins == ["base", "a1", "a2"]
    np.testing.assert_allclose(h2.read_gw_annotation_masses(out), annot.sum(axis=0))
    # heritability 0.3 over all SNPs plus 0.2 on the first annotation, none on the second
    sim = simulate.gwas(g, annot, [0.3, 0.2, 0.0], seed=5)
    g.close()
    y = h2.h2_response(sim.beta, sim.se, sim.n, sim.n)
    fit = h2.fit_h2(L, annot, y, h2.effective_n_scale(sim.n), h2.JackknifeDesign.blocks(M, 100))
    print(f"h2 {fit.h2:.4f} +- {fit.h2_se:.4f} (simulated {sim.h2_total:.4f}); "
          f"a1 {fit.h2_cat[1]:.4f} +- {fit.h2_cat_se[1]:.4f}", flush=True)
    assert 0.0 < fit.h2_se < 0.2
    assert abs(fit.h2 - 0.5) < 4.0 * fit.h2_se
    truth_a1 = 0.2 + 0.3 * annot[:, 1].sum() / M
    assert abs(fit.h2_cat[1] - truth_a1) < 4.0 * fit.h2_cat_se[1]
    assert fit.enrichment[1] > 1.0
