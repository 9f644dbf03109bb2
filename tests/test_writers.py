"""LD-score table and annotation-mass writers: the reference's text formats
(src/summit/ldscore/gw_ldscore.py:1186-1208)."""
from __future__ import annotations

import gzip
import os
import tempfile

import numpy as np
import pandas as pd


def test_ldscore_table_format():
    from summit_amd.ldscore import write_gw_annotation_masses, write_gw_ldscore_table
    d = tempfile.mkdtemp()
    snp = pd.DataFrame({"CHR": [1, 1, 2], "SNP": ["rs1", "rs2", "rs3"], "BP": [100, 250, 77]})
    scores = pd.DataFrame({"base": [1.25, 2.0, -0.5], "annot1": [0.1234567, 0.0, 3.0]})
    path = write_gw_ldscore_table(os.path.join(d, "out"), snp, scores, snp_var=np.array([0.5, 0.25, 1.0]),
                                  ld_rel=np.array([[0.01, 0.02], [0.03, 0.04], [0.05, 0.06]]))
    assert path.endswith(".gw.ldscore.gz")
    with gzip.open(path, "rt") as f:
        lines = f.read().splitlines()
    assert lines[0].split("\t") == ["CHR", "SNP", "BP", "V", "base", "annot1", "LD_REL_base", "LD_REL_annot1"]
    assert lines[1].split("\t") == ["1", "rs1", "100", "0.500000", "1.250000", "0.123457", "0.010000", "0.020000"]
    assert lines[3].split("\t")[4] == "-0.500000"
    mpath = write_gw_annotation_masses(os.path.join(d, "out"), [10000.0, 1234.5678901234, 3])
    assert open(mpath).read() == "10000\t1234.56789\t3\n"


def test_delta_matrix_format():
    """<out>.gw.delta: the B x B delta matrix with the annotation columns as row and
    column names, tab-separated, %.8e (src/summit/ldscore/gw_ldscore.py:2343-2349)."""
    from summit_amd.ldscore import write_gw_delta
    d = tempfile.mkdtemp()
    path = write_gw_delta(os.path.join(d, "out"), np.array([[1.5, -2.0e-3], [0.0, 12345.678]]), ["L2_0", "L2_1"])
    assert path.endswith(".gw.delta")
    lines = open(path).read().splitlines()
    assert lines[0] == "\tL2_0\tL2_1"
    assert lines[1] == "L2_0\t1.50000000e+00\t-2.00000000e-03"
    assert lines[2] == "L2_1\t0.00000000e+00\t1.23456780e+04"
