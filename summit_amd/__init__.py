"""summit_amd — MI355X-native genotype-probe engine for SUMMIT-style stochastic LD scores.

Hot path: X.Z and X^T(X.Z) over packed 2-bit PLINK BED genotypes resident in
HBM, computed by hand-written HIP kernels for gfx950 (libsummit_gpe.so).

* ``summit_amd.engine``   — device-level API (Genotype handle, phase kernels)
* ``summit_amd.gwldcore`` — drop-in mirror of the reference's ``summit.gwldcore``
* ``summit_amd.ldscore``  — genome-wide / partitioned LD-score driver
"""
__version__ = "0.1.0"
