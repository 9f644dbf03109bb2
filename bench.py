"""Benchmark of the genome-wide stochastic LD-score hot path on one MI355X.

Workload (the headline configuration): N = 100,000 samples, M = 500,000 SNPs,
B = 64 Rademacher probes, packed 2-bit genotypes from the block-Markov
generator (two haplotypes per sample, 100-SNP blocks, copy probability 0.9,
MAF ~ U(0.01, 0.5), 0.5% missing calls), mean imputation, one annotation bin,
block size 1000.  The genotype matrix is generated on the GPU and stays
resident in HBM.

One step = one probe tile of the estimator: phase 1 (X.Z over all M SNPs,
probe streams drawn per SNP block on the GPU) + phase 2 (X^T(X.Z) with the
squared-norm reduction into the per-SNP accumulators).  The metric counts
genotype-probe pairs, 2*N*M*B per step (both traversals).

Prints one JSON line (rank 0).  --gpus N > 1 is refused: this engine runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0  # dense int8 MFMA, 5.0 POP/s


def kernel_source_sha() -> str:
    """SHA-256 of the kernel sources (the contraction kernels and their launch
    code): profiles/traffic.json carries the hash of the build its counters
    were read from, and a value from any other build is not reported."""
    import hashlib
    h = hashlib.sha256()
    for name in ("gpe_kernels.hip", "gpe_api.hip", "gpe_internal.h"):
        with open(os.path.join(ROOT, "summit_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(geno, n_snps_sample: int, b: int, seed: int, step: int = 1000, repeats: int = 3) -> dict:
    """The reference's CPU dense path on a bounded sample of the same workload,
    in float32 (the reference's default dtype, src/summit/cli.py:796) and in
    float64: per SNP block the reference's own in-place standardized decode
    (genotype.cpp through oracle/_ref when built, else the oracle's
    restatement), the residual scale, the block's probes, the phase-1 GEMM
    Xz += X (inv*Z); then per block a second decode, the phase-2 GEMM
    Work = X^T Xz and the squared-norm reduce in float64 -- GEMMs through
    NumPy's OpenBLAS on every host core.  One warm-up block, then the best of
    `repeats` full passes per precision."""
    import ctypes
    import tempfile

    from oracle import oracle as o

    n = geno.n
    L = n_snps_sample
    body = geno.download_rows("snp_raw", 0, L)
    cores = len(os.sched_getaffinity(0))
    ref_path = os.path.join(ROOT, "oracle", "_ref", "libref_genotype.so")
    tmp = tempfile.mkdtemp(prefix="gpe_cpu_")
    prefix = os.path.join(tmp, "sample")
    with open(prefix + ".bed", "wb") as f:
        f.write(o.BED_MAGIC)
        f.write(body.tobytes())
    with open(prefix + ".fam", "w") as f:
        f.write("x\n" * n)
    with open(prefix + ".bim", "w") as f:
        f.write("x\n" * L)
    use_ref = os.path.exists(ref_path)
    if use_ref:
        ref = ctypes.CDLL(ref_path)
        ref.ref_read_block_standardized_into.argtypes = [
            ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
            ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        rows = np.arange(n, dtype=np.int32)
    bed_b, fam_b = (prefix + ".bed").encode(), (prefix + ".fam").encode()

    def read_block(s, e, dt, buf):
        if use_ref:
            rc = ref.ref_read_block_standardized_into(bed_b, fam_b, s, e, rows.ctypes.data, n, 1, 0, 0,
                                                      buf.ctypes.data, buf.size, int(dt == np.float32))
            if rc != 0:
                raise RuntimeError("reference decoder failed")
            return buf[:(e - s) * n].reshape(e - s, n).T
        return o.read_block_standardized(body, n, s, e, None, 1, "mean", 0).astype(dt)

    blocks = [(s, min(L, s + step)) for s in range(0, L, step)]
    probes = {s: o.probe_block(seed, s, 0, e - s, b) for s, e in blocks}

    def one_pass(dt, blks):
        buf = np.empty(step * n, dtype=dt)
        xz = np.zeros((n, b), dtype=dt, order="F")
        invs = {}
        for s, e in blks:
            X = read_block(s, e, dt, buf)
            ss = np.einsum("ij,ij->j", X, X, dtype=np.float64)
            inv = 1.0 / np.sqrt(np.maximum(ss / (n - 1), 1e-10))
            invs[s] = inv
            xz += X @ (inv[:, None] * probes[s]).astype(dt)
        meansq = np.zeros(L)
        for s, e in blks:
            X = read_block(s, e, dt, buf)
            W = X.T @ xz
            meansq[s:e] += np.einsum("ij,ij->i", W, W, dtype=np.float64) * invs[s] ** 2 / (n - 1.0) ** 2
        return meansq

    out = {}
    t_all = time.perf_counter()
    for name, dt in (("float32", np.float32), ("float64", np.float64)):
        one_pass(dt, blocks[:1])  # warm-up: decoder, BLAS threads, page cache
        best = np.inf
        for _ in range(repeats):
            t0 = time.perf_counter()
            one_pass(dt, blocks)
            best = min(best, time.perf_counter() - t0)
        out[name] = {"gop_s": round(2.0 * n * L * b / best / 1e9, 3), "best_s": round(best, 3)}
    return {"value": out["float32"]["gop_s"], "unit": "Gop/s", "cores": cores,
            "kind": "reference" if use_ref else "port",
            "float32": out["float32"], "float64": out["float64"],
            "sample": f"N={n} x first {L} SNPs of the same genotypes, B={b}, both passes "
                      f"({'reference genotype.cpp in-place decode' if use_ref else 'oracle decode'} + OpenBLAS GEMMs), "
                      f"value = float32 (the reference default); best of {repeats} after a warm-up block, "
                      f"{time.perf_counter() - t_all:.1f} s in all"}


def wall_to_one_percent(engine, ldscore, torch, dev) -> dict:
    """Wall time for the plain GPU estimator to reach a mean relative LD-score
    error <= 1% at N=5,000, M=10,000 (i.i.d. genotypes, MAF ~ U(0.05, 0.5), 1%
    missing, seeds 101 / 1101).  The exact target L_j = sum_i r_ij^2 - M/(N-1)
    (r with the residual scales of the estimator) comes from a dense FP64
    X X^T on the GPU.  Probe counts double from 16,384 until the error bar is
    met; the reported time is genotype preparation plus the run that met it.
    V-tiles of 256 probes, eight thousand probes per traversal pair."""
    N, M, root = 5000, 10000, 1101
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = engine.Genotype.synthetic(N, M, seed=101, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.0)
    g.prepare(ddof=1, impute="mean", impute_seed=ldscore.impute_seed_from_root(root))
    torch.cuda.synchronize()
    t_prep = time.perf_counter() - t0
    # exact target from the raw codes and the engine's own standardization parameters
    st = g.stats()
    codes = torch.from_numpy(g.download("snp_raw")).to(dev)
    sh = torch.tensor([0, 2, 4, 6], device=dev, dtype=torch.uint8)
    c = ((codes[:, :, None] >> sh) & 3).reshape(M, -1)[:, :N]
    val = torch.tensor([0.0, 0.0, 1.0, 2.0], device=dev, dtype=torch.float64)[c.long()]
    obs = (c != 1).to(torch.float64)
    mu = torch.from_numpy(st["mean"]).to(dev)[:, None]
    isd = torch.from_numpy(st["inv_sd"]).to(dev)[:, None]
    x = (val - mu) * isd * obs
    inv = engine.resvar_inv(g, 0, M)
    xs = x * inv[:, None]
    r = (xs @ xs.T) / float(N - 1)
    exact = ((r * r).sum(dim=1) - M / float(N - 1)).cpu().numpy()
    del codes, c, val, obs, x, xs, r
    tried = []
    out = None
    for B in (16384, 32768, 65536, 131072, 262144):
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        # V-tiles of 256 probes (the reference's own granularity rule), served 8,192 at a time
        res = ldscore.genomewide_ldscore(g, nvecs=B, seed=root, vtiles=ldscore.bounded_vtile_sizes(B, 256),
                                         traversal_width=8192)
        t_run = time.perf_counter() - t1
        err = float(np.mean(np.abs(res.ldscore[:, 0] - exact) / np.maximum(np.abs(exact), 1e-12)))
        tried.append({"probes": B, "seconds": round(t_run, 4), "mean_rel_err": round(err, 5)})
        log(f"C1 plain estimator B={B}: {t_run:.3f} s, mean rel err {err:.4f}")
        if err <= 0.01:
            out = {"probes": B, "seconds": round(t_prep + t_run, 4), "estimator_seconds": round(t_run, 4),
                   "mean_rel_err": round(err, 5)}
            break
    g.close()
    return {"workload": "N=5000 M=10000 i.i.d. genotypes, 1% missing, Rademacher, plain estimator, "
                        "exact dense FP64 target", "reached": out, "tried": tried}


def windowed_rate(engine, torch) -> dict:
    """Exact windowed LD scores, N=50,000, M=300,000 on 22 chromosomes (block-Markov
    genotypes, seed 404), SNP positions from a Poisson process with mean spacing
    3 kb, +-1 Mb windows: every chromosome's windows in one banded-Gram launch,
    against the baseline of one launch per chromosome.  Pairs = N x sum of
    window sizes (each SNP with every SNP of its window, itself included)."""
    from summit_amd import winldcore
    N, M, n_chr = 50_000, 300_000, 22
    g = engine.Genotype.synthetic(N, M, seed=404, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    rng = np.random.default_rng(1404)
    bounds = np.linspace(0, M, n_chr + 1).astype(np.int64)
    chrs = []
    pairs = 0
    for c in range(n_chr):
        s0, s1 = int(bounds[c]), int(bounds[c + 1])
        bp = np.cumsum(rng.exponential(3000.0, s1 - s0)).astype(np.int64) + 1
        lo = np.searchsorted(bp, bp - 1_000_000, side="left")
        hi = np.searchsorted(bp, bp + 1_000_000, side="right")
        pairs += int((hi - lo).sum()) * N
        chrs.append((s0, s1, bp))
    winldcore.windowed_ld(g, chrs[0][0], chrs[0][1], bp=chrs[0][2], ld_wind_kb=1000.0)  # warm-up
    winldcore.windowed_ld_multi(g, chrs, ld_wind_kb=1000.0)
    # baseline: one launch per chromosome
    torch.cuda.synchronize()
    engine.profile_enable(True)
    engine.profile_reset()
    t0 = time.perf_counter()
    per_chr = [winldcore.windowed_ld(g, s0, s1, bp=bp, ld_wind_kb=1000.0) for s0, s1, bp in chrs]
    torch.cuda.synchronize()
    t_chr = time.perf_counter() - t0
    gram_chr_ms, gram_chr_n = engine.profile_read("gram")
    # every chromosome's windows in one launch
    engine.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    allc = winldcore.windowed_ld_multi(g, chrs, ld_wind_kb=1000.0)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    gram_ms, gram_n = engine.profile_read("gram")
    engine.profile_enable(False)
    g.close()
    ref = np.concatenate(per_chr, axis=0)
    return {"workload": "N=50000 M=300000, 22 chromosomes, Poisson positions (3 kb), +-1 Mb windows",
            "seconds": round(t, 4), "pairs": pairs, "gpairs_per_s": round(pairs / t / 1e9, 1),
            "gram_kernel_ms": round(gram_ms, 3), "gram_launches": gram_n,
            "per_chromosome_launches": {"seconds": round(t_chr, 4), "gram_kernel_ms": round(gram_chr_ms, 3),
                                        "gram_launches": gram_chr_n},
            "max_abs_diff_vs_per_chromosome": float(np.max(np.abs(allc - ref)))}


def gxe_rate(engine, torch) -> dict:
    """SUMMIT-GxE directional scores at N=100,000, M=500,000, B=64 for K=8
    environments (6 standardised N(0,1), 2 Bernoulli(0.3)), block-Markov genotypes
    (seed 303), environments and probes from root seed 1303.  Pairs per
    environment: 2B sources (N*M*2B) plus two target traversals (N*M*4B)."""
    from summit_amd import gxe
    N, M, B, K = 100_000, 500_000, 64, 8
    # warm-up on a small matrix: first launches of every kernel the path uses
    # (module loads are a once-per-process cost, not part of the workload)
    gw = engine.Genotype.synthetic(2000, 3000, seed=1, maf_lo=0.05, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                   copy_prob=0.9)
    gw.prepare(ddof=1, impute="mean")
    ew = np.random.default_rng(0).standard_normal((2000, K))
    gxe.gxe_ldscore_multi(gw, (ew - ew.mean(axis=0)) / ew.std(axis=0), None, nvecs=B, seed=1)
    gw.close()
    g = engine.Genotype.synthetic(N, M, seed=303, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    rng = np.random.default_rng(1303)
    envs = np.empty((N, K))
    for k in range(6):
        e = rng.standard_normal(N)
        envs[:, k] = (e - e.mean()) / e.std()
    envs[:, 6:] = (rng.random((N, 2)) < 0.3).astype(np.float64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    gxe.gxe_ldscore_multi(g, envs, None, nvecs=B, seed=1303)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    g.close()
    pairs = 6 * K * B * N * M
    return {"workload": "N=100000 M=500000 B=64 K=8 environments, no covariates", "seconds": round(t, 3),
            "pairs": pairs, "gpairs_per_s": round(pairs / t / 1e9, 1)}


def cpu_mailman_c1(engine, torch) -> dict:
    """The reference's CPU path for at most 10 probes (Mailman lookup tables,
    OpenMP over segments; the oracle's restatement, pinned to the dense integer
    products) on N=5,000, M=10,000 i.i.d. genotypes with 1% missing calls and
    B=10 Rademacher probes: G.Z then G^T (G.Z), exact int64, on all host cores,
    beside the GPU's exact integer products on the same genotypes."""
    import numpy as np

    from oracle import oracle as o

    n, m, b = 5000, 10_000, 10
    g = engine.Genotype.synthetic(n, m, seed=101, maf_lo=0.05, maf_hi=0.5, miss_rate=0.01, block_len=100,
                                  copy_prob=0.0)
    g.prepare(ddof=1, impute="mean")
    packed = np.ascontiguousarray(g.download_rows("snp_raw", 0, m)[:, :(n + 3) // 4])
    cores = len(os.sched_getaffinity(0))
    rng = np.random.default_rng(1101)
    z = (rng.integers(0, 2, size=(m, b)) * 2 - 1).astype(np.int8)
    u = o.mailman_gz(packed, n, z, threads=cores)  # warm-up
    reps, t0 = 0, time.perf_counter()
    while reps < 3 or time.perf_counter() - t0 < 2.0:
        u = o.mailman_gz(packed, n, z, threads=cores)
        w = o.mailman_gtu(packed, n, u, threads=cores)
        reps += 1
    cpu_s = (time.perf_counter() - t0) / reps
    pairs = 2.0 * n * m * b
    zd = torch.from_numpy(z).to(engine.default_device())
    ud = engine.gz_i32(g, zd)
    wd = engine.gtu_i64(g, ud)
    torch.cuda.synchronize()
    same = bool(np.array_equal(ud.cpu().numpy().T.astype(np.int64), u)) and bool(np.array_equal(wd.cpu().numpy(), w))
    t1 = time.perf_counter()
    for _ in range(20):
        ud = engine.gz_i32(g, zd)
        wd = engine.gtu_i64(g, ud)
    torch.cuda.synchronize()
    gpu_s = (time.perf_counter() - t1) / 20
    g.close()
    return {"workload": f"N={n} M={m} B={b} Rademacher, i.i.d. genotypes, 1% missing (as 0): G.Z and G^T(G.Z)",
            "cpu_mailman": {"gpairs_per_s": round(pairs / cpu_s / 1e9, 2), "seconds": round(cpu_s, 4),
                            "cores": cores, "reps": reps, "kind": "port",
                            "segment_snps": o.mailman_segment_size(n)},
            "gpu_exact_int": {"gpairs_per_s": round(pairs / gpu_s / 1e9, 1), "seconds": round(gpu_s, 6)},
            "results_equal": same}


def ingest_rate(engine, torch, geno, n_snps: int) -> dict:
    """Host -> HBM rate of the BED path: the first n_snps packed SNP rows of the
    benchmark genotypes, copied to host memory, then uploaded (pinned,
    double-buffered staging) and prepared (statistics, transpose, missing lists)."""
    body = geno.download_rows("snp_raw", 0, n_snps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g = engine.Genotype.from_bed_bytes(body, geno.n, n_snps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    g.prepare(ddof=1, impute="mean")
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    g.close()
    return {"bytes": int(body.nbytes), "upload_s": round(t1 - t0, 4), "prepare_s": round(t2 - t1, 4),
            "upload_gbs": round(body.nbytes / (t1 - t0) / 1e9, 2),
            "upload_prepare_gbs": round(body.nbytes / (t2 - t0) / 1e9, 2)}


def partitioned(engine, ldscore, torch) -> dict:
    """Partitioned LD scores, N=50,000, M=200,000 (block-Markov genotypes, seed
    202, MAF ~ U(0.01, 0.5), 0.5% missing), 8 binary annotations each a random
    10% of SNPs, B=100 Rademacher probes (root 1202), Monte Carlo variances on:
    the 5-limb and 6-limb int8 paths (their difference bounds the fixed-point
    error) and the 5-limb error against the exact partitioned LD scores
    (full-window Gram) in Monte Carlo standard errors.  Pairs per run: N*M*B*K
    in each traversal (every bin's columns run over all SNPs)."""
    from summit_amd import cv
    N, M, B, K = 50_000, 200_000, 100, 8
    g = engine.Genotype.synthetic(N, M, seed=202, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    annot = (np.random.default_rng(1202).random((M, K)) < 0.1).astype(np.float64)
    out = {"workload": "N=50000 M=200000, 8 random 10% annotations, B=100 Rademacher, mean imputation"}
    res = {}
    for limbs in (5, 6):
        ldscore.genomewide_ldscore(g, B, seed=1, annot=annot, limbs=limbs, mc=True)  # warm-up, same shapes
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res[limbs] = ldscore.genomewide_ldscore(g, B, seed=1202, annot=annot, limbs=limbs, mc=True)
        torch.cuda.synchronize()
        out[f"seconds_{limbs}_limbs"] = round(time.perf_counter() - t0, 4)
    pairs = 2.0 * N * M * B * K
    out["pairs"] = pairs
    out["gpairs_per_s_5_limbs"] = round(pairs / out["seconds_5_limbs"] / 1e9, 1)
    l5, l6 = res[5].ldscore, res[6].ldscore
    # relative to each bin's mean absolute score (single partitioned scores can be ~0)
    out["max_rel_diff_5_vs_6_limbs"] = float(np.max(np.abs(l5 - l6) / np.mean(np.abs(l6), axis=0)[None, :]))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    exact = cv.exact_ldscore(g, annot)
    out["exact_s"] = round(time.perf_counter() - t0, 4)
    g.close()
    # statistical check of the reference's own tests: |error| within 5 Monte Carlo standard errors
    se = np.sqrt(np.maximum(res[5].mc_var, 0.0))
    zs = np.abs(l5 - exact) / np.maximum(se, 1e-300)
    out["frac_within_5_mc_se"] = round(float(np.mean(zs <= 5.0)), 6)
    out["mean_abs_z"] = round(float(np.mean(zs)), 4)
    out["mean_abs_err_vs_exact"] = round(float(np.mean(np.abs(l5 - exact))), 5)
    return out


def dropin_loop(engine, ldscore, torch) -> dict:
    """What a caller of the drop-in module gets: the reference driver's per-block
    loop (src/summit/ldscore/gw_ldscore.py:1905-2133) through summit_amd.gwldcore
    with NumPy arrays, at N=50,000, M=200,000 (block-Markov genotypes, seed 202,
    0.5% missing, mean imputation), one bin, B=100 Rademacher probes in one
    V-tile, blocks of 1,000 SNPs, the residual scales fused into phase 1; the BED
    is written to a temporary file and made resident by the first call (timed
    separately).  Beside it the engine's fast path (one call per phase over all
    blocks) on the same genotypes."""
    import tempfile

    from oracle import oracle as o
    from summit_amd import gwldcore as gw
    N, M, B, step, seed = 50_000, 200_000, 100, 1000, 1202
    g = engine.Genotype.synthetic(N, M, seed=202, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    body = g.download_rows("snp_raw", 0, M)
    with tempfile.TemporaryDirectory() as d:
        prefix = os.path.join(d, "c2")
        with open(prefix + ".bed", "wb") as f:
            f.write(o.BED_MAGIC)
            f.write(body.tobytes())
        with open(prefix + ".fam", "w") as f:
            f.write("x\n" * N)
        with open(prefix + ".bim", "w") as f:
            f.write("x\n" * M)
        del body
        kw = dict(bed_prefix=prefix, fam_path=prefix + ".fam", row_sel=None, ddof=1, impute_mode="mean",
                  impute_seed=None)
        t0 = time.perf_counter()
        gw.prefetch_bed_block(prefix, prefix + ".fam", 0, step, 1, "mean", 1, None)
        t_resident = time.perf_counter() - t0
        blocks = [(s, min(M, s + step)) for s in range(0, M, step)]
        annot = np.ones((M, 1))

        def run():
            inv_all = np.empty(M)
            meansq = np.zeros((M, 1))
            xz = np.empty((N, B), order="F")
            init = np.zeros(1, dtype=np.uint8)
            for s, e in blocks:
                gw.phase1_compute_Xz_bed_chunk(blk_start=s, blk_end=e, annot_blk=annot[s:e], inv_right=inv_all[s:e],
                                               v_start=0, v_count=B, kmax_hint=e - s, seed=seed, Xz2d_chunk=xz,
                                               bin_init_mask=init, inv_out=inv_all[s:e], **kw)
            for s, e in blocks:
                gw.phase2_accum_XtXz_bed(blk_start=s, blk_end=e, inv_left=inv_all[s:e], tile_nvecs=B, Xz2d=xz,
                                         meansq_accum=meansq, N_denom=N, **kw)
            return meansq

        run()  # warm-up: workspaces, schedules
        t0 = time.perf_counter()
        ms_dropin = run()
        t_dropin = time.perf_counter() - t0
        gw._clear_bed_mapping_cache()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fast = ldscore.genomewide_ldscore(g, B, seed=seed, annot=annot)
    t_fast = time.perf_counter() - t0
    g.close()
    pairs = 2.0 * N * M * B
    return {"workload": "N=50000 M=200000, one bin, B=100, 200 blocks of 1000 SNPs, NumPy arrays through "
                        "summit_amd.gwldcore (the reference driver's loop)",
            "dropin_s": round(t_dropin, 3), "dropin_gop_s": round(pairs / t_dropin / 1e9, 1),
            "fast_path_s": round(t_fast, 3), "fast_path_gop_s": round(pairs / t_fast / 1e9, 1),
            "bed_resident_s": round(t_resident, 3),
            "max_rel_diff_vs_fast_path": float(np.max(np.abs(ms_dropin[:, 0] / B - (fast.ldscore[:, 0] + M / (N - 1.0)))
                                                      / np.abs(fast.ldscore[:, 0] + M / (N - 1.0))))}


def gxe_direct_loop(engine, torch) -> dict:
    """What a caller of the GxE drop-in gets: the reference driver's direct
    per-block loop (src/summit/ldscore/gwe_ldscore.py:2218-2260, :2716-2830,
    :3683-3759) through summit_amd.gxeldcore.DirectContext with NumPy arrays, at
    N=50,000, M=100,000 (block-Markov genotypes, seed 404, no missing calls, as the
    production gate requires), one environment (standard normal, root 1404), the
    design basis [1, e], one bin, B=16 Rademacher probes drawn by NumPy Philox per
    (block, probe) as the driver draws them, blocks of 1,000 SNPs: features per
    block, sources per block summed into the global sketch, the sketch sealed, then
    targets per block with Work^2 accumulated into XX / XW / WX / WW.  The BED is
    written to a temporary file and made resident when the context is built (timed
    separately).  Beside it the engine's fast path (gxe.gxe_ldscore, one traversal
    per phase over all blocks) on the same genotypes and probes.  Multiply-adds:
    sources N*M*2B, targets N*M*4B."""
    import tempfile

    from oracle import oracle as o
    from summit_amd import gxe, gxeldcore
    N, M, B, step, root = 50_000, 100_000, 16, 1000, 1404
    g = engine.Genotype.synthetic(N, M, seed=404, maf_lo=0.05, maf_hi=0.5, miss_rate=0.0, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    body = g.download_rows("snp_raw", 0, M)
    rng = np.random.default_rng(root)
    env = rng.standard_normal(N)
    env = (env - env.mean()) / env.std(ddof=1)
    q = np.asfortranarray(np.linalg.qr(np.column_stack([np.ones(N), env]))[0])
    blocks = [(s, min(M, s + step)) for s in range(0, M, step)]
    df = float(N - 2)
    with tempfile.TemporaryDirectory() as d:
        prefix = os.path.join(d, "gxe")
        with open(prefix + ".bed", "wb") as f:
            f.write(o.BED_MAGIC)
            f.write(body.tobytes())
        with open(prefix + ".fam", "w") as f:
            f.write("x\n" * N)
        with open(prefix + ".bim", "w") as f:
            f.write("x\n" * M)
        del body
        fds = [os.open(prefix + ext, os.O_RDONLY | os.O_CLOEXEC) for ext in (".bed", ".bim", ".fam")]
        try:
            def run(ctx):
                sx_all, sw_all = np.empty(M), np.empty(M)
                for s, e in blocks:
                    f = ctx.feature_block(blk_start=s, blk_end=e, eps_var=1e-10, require_missing_free=False)
                    sx_all[s:e], sw_all[s:e] = f["scale_x"], f["scale_w"]
                sketch = np.zeros((N, 2 * B), order="F")
                for s, e in blocks:
                    z = np.empty((e - s, B), order="F")
                    for v in range(B):
                        gen = np.random.Generator(np.random.Philox(gxe.make_seed(root, s, v)))
                        z[:, v] = 2.0 * gen.integers(0, 2, size=e - s, dtype=np.int8) - 1.0
                    sx_, sw_, _ = ctx.source_block(s, e, sx_all[s:e], sw_all[s:e], np.ones(e - s), z, True)
                    sketch[:, :B] += sx_
                    sketch[:, B:] += sw_
                panel = ctx.prepare_projected_sources(sketch, 1e-9)
                acc = np.zeros((M, 4))
                for s, e in blocks:
                    wx, ww, _, _ = ctx.target_projected_block(s, e, sx_all[s:e], sw_all[s:e], panel, True)
                    acc[s:e, 0] = np.sum(wx[:, :B] ** 2, axis=1)
                    acc[s:e, 1] = np.sum(wx[:, B:] ** 2, axis=1)
                    acc[s:e, 2] = np.sum(ww[:, :B] ** 2, axis=1)
                    acc[s:e, 3] = np.sum(ww[:, B:] ** 2, axis=1)
                return acc / (df * df * B)

            warm = gxeldcore.DirectContext(fds[0], fds[1], fds[2], None, 1, env, q, 1, 64 << 30, 64)
            run(warm)  # warm-up: workspaces, schedules
            warm.close()
            # a fresh context for the timed loop: its first feature call computes every SNP's features
            t0 = time.perf_counter()
            ctx = gxeldcore.DirectContext(fds[0], fds[1], fds[2], None, 1, env, q, 1, 64 << 30, 64)
            t_resident = time.perf_counter() - t0
            t0 = time.perf_counter()
            scores = run(ctx)
            t_loop = time.perf_counter() - t0
            ctx.close()
        finally:
            for fd in fds:
                os.close(fd)
    c_int = (env / np.linalg.norm(env))[:, None]
    gxe.gxe_ldscore(g, env, C_int=c_int, nvecs=B, seed=root, step_size=step, limbs=6)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fast = gxe.gxe_ldscore(g, env, C_int=c_int, nvecs=B, seed=root, step_size=step, limbs=6)
    t_fast = time.perf_counter() - t0
    g.close()
    ref = np.column_stack([fast[k][:, 0] for k in ("xx", "xw", "wx", "ww")])
    macs = 6.0 * N * M * B
    return {"workload": "N=50000 M=100000, one environment, basis [1, e], one bin, B=16, 100 blocks of 1000 SNPs, "
                        "NumPy arrays through summit_amd.gxeldcore.DirectContext (the reference driver's direct loop)",
            "dropin_s": round(t_loop, 3), "dropin_gmac_s": round(macs / t_loop / 1e9, 1),
            "fast_path_s": round(t_fast, 3), "fast_path_gmac_s": round(macs / t_fast / 1e9, 1),
            "bed_resident_s": round(t_resident, 3),
            "max_rel_diff_vs_fast_path": float(np.max(np.abs(scores - ref) / np.abs(ref)))}


def h2_end_to_end(engine, ldscore, torch) -> dict:
    """The LD scores' consumer end to end at N=50,000, M=200,000 (block-Markov
    genotypes, seed 202, 0.5% missing): partitioned LD scores (a base bin plus 8
    random 10% annotations, B=100 Rademacher probes) written to and read back from
    .gw.ldscore.gz / .gw.M, a GWAS simulated on the same genotypes with
    heritability 0.3 spread over all SNPs plus 0.2 on the first annotation, and
    the moment solve with a 200-block delete-one jackknife."""
    import tempfile

    import pandas as pd

    from summit_amd import h2, simulate
    N, M, B = 50_000, 200_000, 100
    t_all = time.perf_counter()
    g = engine.Genotype.synthetic(N, M, seed=202, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                  copy_prob=0.9)
    g.prepare(ddof=1, impute="mean")
    annot = np.column_stack([np.ones(M), (np.random.default_rng(1202).random((M, 8)) < 0.1).astype(np.float64)])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = ldscore.genomewide_ldscore(g, B, seed=1202, annot=annot)
    t_ld = time.perf_counter() - t0
    names = ["base"] + [f"a{k}" for k in range(1, 9)]
    with tempfile.TemporaryDirectory() as d:
        out = os.path.join(d, "c2")
        t0 = time.perf_counter()
        snp = pd.DataFrame({"CHR": np.ones(M, dtype=np.int64), "SNP": [f"rs{j}" for j in range(M)],
                            "BP": np.arange(M, dtype=np.int64) * 100})
        ldscore.write_gw_ldscore_table(out, snp, pd.DataFrame(res.ldscore, columns=names))
        ldscore.write_gw_annotation_masses(out, annot.sum(axis=0))
        _, _, L, bins = h2.read_gw_ldscores(out)
        t_io = time.perf_counter() - t0
    h2_bins = np.zeros(9)
    h2_bins[0], h2_bins[1] = 0.3, 0.2
    t0 = time.perf_counter()
    sim = simulate.gwas(g, annot, h2_bins, seed=505)
    t_gwas = time.perf_counter() - t0
    g.close()
    t0 = time.perf_counter()
    y = h2.h2_response(sim.beta, sim.se, sim.n, sim.n)
    fit = h2.fit_h2(L, annot, y, h2.effective_n_scale(sim.n), h2.JackknifeDesign.blocks(M, 200))
    t_fit = time.perf_counter() - t0
    truth_a1 = 0.2 + 0.3 * annot[:, 1].sum() / M
    return {"workload": "N=50000 M=200000, base + 8 random 10% annotations, B=100, 200 jackknife blocks",
            "h2": round(fit.h2, 5), "h2_se": round(fit.h2_se, 5), "h2_true": 0.5,
            "h2_simulated": round(sim.h2_total, 5), "z_h2": round((fit.h2 - 0.5) / fit.h2_se, 3),
            "h2_a1": round(float(fit.h2_cat[1]), 5), "h2_a1_se": round(float(fit.h2_cat_se[1]), 5),
            "h2_a1_true": round(truth_a1, 5), "enrichment_a1": round(float(fit.enrichment[1]), 4),
            "bins_read": len(bins), "ldscore_s": round(t_ld, 3), "files_s": round(t_io, 3),
            "gwas_s": round(t_gwas, 3), "fit_s": round(t_fit, 3), "total_s": round(time.perf_counter() - t_all, 3)}


def vendor_int8_gemm(torch, dev, rows: int = 131072, k: int = 102400, n: int = 320) -> dict:
    """The ROCm library int8 GEMM (torch._int_mm, int32 out) on the limb product of the phase-2
    contraction with its operands already unpacked: SNP rows x samples (K = N rounded to a multiple
    of 8) times samples x (5 limbs x 64 columns), on a 131,072-row slice (the rate is flat in the row
    count).  It does the same MFMA work per row as the contraction without the 2-bit unpack, the
    missing-call terms or the fused epilogue; the square 16384^3 int8 GEMM is its best case."""
    out = {}
    for name, (m_, k_, n_) in (("phase2_limb_shape", (rows, k, n)), ("square_16384", (16384, 16384, 16384))):
        a = torch.randint(-128, 128, (m_, k_), dtype=torch.int8, device=dev)
        b = torch.randint(-128, 128, (k_, n_), dtype=torch.int8, device=dev)
        torch._int_mm(a, b)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            torch._int_mm(a, b)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1000.0 / 5
        out[name] = {"m": m_, "k": k_, "n": n_, "ms": round(ms, 3),
                     "tops": round(2.0 * m_ * k_ * n_ / (ms / 1000.0) / 1e12, 1)}
        del a, b
    return out


def learned_cv(engine, ldscore, torch, dev, plain_max: int = 98304, plain_tile: int = 2048) -> dict:
    """Learned control variate at N=50,000, M=200,000 (block-Markov genotypes,
    MAF ~ U(0.01, 0.5), 0.5% missing, 8 random 10% annotations).  The network
    is trained on seed 505 against a 1,000-probe plain run (root 1505) and
    evaluated on seed 506 (root 1506): control variate with B=8 probes vs plain
    B=8, plain B=1,000 and the exact LD scores (full-window int8 Gram).  The
    plain estimator's time to <= 1% mean relative error comes from one run in
    tiles of ``plain_tile`` probes, the error of the running mean checked after
    each tile outside the timed region."""
    from summit_amd import cv, mlp
    N, M = 50_000, 200_000

    def data(seed):
        g = engine.Genotype.synthetic(N, M, seed=seed, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                      copy_prob=0.9)
        g.prepare(ddof=1, impute="mean")
        annot = (np.random.default_rng(seed + 1000).random((M, 8)) < 0.1).astype(np.float64)
        return g, annot

    out = {"workload": "N=50000 M=200000 block-Markov genotypes, 8 annotations; train seed 505 (1000-probe "
                       "labels), evaluate seed 506; error = mean |L_hat - L| / L against exact LD"}
    ga, annot_a = data(505)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fa = cv.snp_features(ga, annot_a)
    t1 = time.perf_counter()
    lab = ldscore.genomewide_ldscore(ga, 1000, seed=1505, vtiles=ldscore.bounded_vtile_sizes(1000, 512), mc=True)
    t2 = time.perf_counter()
    model = cv.fit(fa, lab.ldscore[:, 0], lab.mc_var[:, 0], epochs=40, batch=4096, seed=0)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    ga.close()
    out["training"] = {"features_s": round(t1 - t0, 4), "labels_1000_probes_s": round(t2 - t1, 4),
                       "train_s": round(t3 - t2, 4), "epochs": model.info["epochs"],
                       "best_epoch": model.info["best_epoch"], "steps": model.info["steps"],
                       "tau2": model.tau2, "tau2_raw": model.tau2_raw}
    log(f"cv training: {out['training']}")

    gb, annot_b = data(506)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    exact = cv.exact_ldscore(gb)
    out["exact_s"] = round(time.perf_counter() - t0, 4)

    def err(est):
        return float(np.mean(np.abs(est - exact) / np.abs(exact)))

    p1000 = ldscore.genomewide_ldscore(gb, 1000, seed=1506, vtiles=ldscore.bounded_vtile_sizes(1000, 512))
    net = mlp.Net(model.in_dim, max_rows=M)
    cv.cv_ldscore(gb, model, annot_b, nvecs=8, seed=1507, net=net)  # warm-up (kernel loads, workspaces)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = cv.cv_ldscore(gb, model, annot_b, nvecs=8, seed=1506, net=net)
    torch.cuda.synchronize()
    t_cv = time.perf_counter() - t0
    feats_b = cv.snp_features(gb, annot_b)
    base = feats_b[:, 0] + cv.remainder_scale(feats_b[:, 0]) * model.resid_mean  # same features, no network
    # the method without the network: the same combination with the in-window sum plus
    # the mean remainder as the prediction, its squared error calibrated on the training labels
    base_a = fa[:, 0] + cv.remainder_scale(fa[:, 0]) * model.resid_mean
    tau2_base = max(float(np.mean((lab.ldscore[:, 0] - base_a) ** 2) - np.mean(lab.mc_var[:, 0])),
                    (1e-3 * float(np.mean(np.abs(lab.ldscore[:, 0])))) ** 2)
    est_base, _, _ = cv.combine(base, res.plain, tau2_base, 8, M, N)
    out["errors"] = {"cv_b8": round(err(res.ldscore), 5), "plain_b8": round(err(res.plain), 5),
                     "prediction": round(err(res.prediction), 5), "inwindow_plus_mean_resid": round(err(base), 5),
                     "cv_b8_without_network": round(err(est_base), 5),
                     "network_gain_prediction": round(1.0 - err(res.prediction) / err(base), 4),
                     "network_gain_cv_b8": round(1.0 - err(res.ldscore) / err(est_base), 4),
                     "shrinkage_bias_bound": round(res.bias_bound, 5),
                     "plain_b1000": round(err(p1000.ldscore[:, 0]), 5),
                     "cv_b8_vs_plain_b1000": round(float(np.mean(np.abs(res.ldscore - p1000.ldscore[:, 0]) /
                                                                np.abs(p1000.ldscore[:, 0]))), 5)}
    out["cv_seconds"] = round(t_cv, 4)
    out["cv_parts_s"] = {k: round(v, 4) for k, v in res.seconds.items()}
    out["cv_fallback_snps"] = res.fallback
    net.close()
    log(f"cv errors: {out['errors']}, cv {t_cv:.3f} s")

    # plain estimator until <= 1% mean relative error
    inv = engine.resvar_inv(gb, 0, M)
    annot1 = torch.ones((M, 1), dtype=torch.float64, device=dev)
    meansq = torch.zeros((M, 1), dtype=torch.float64, device=dev)
    exact_d = torch.from_numpy(exact).to(dev)
    xz = torch.empty((plain_tile, N), dtype=torch.float64, device=dev)
    elapsed, v0, reached, trace = 0.0, 0, None, []
    while v0 < plain_max:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xz.zero_()
        engine.phase1_xz(gb, 0, M, 1000, annot1, inv, v0, plain_tile, 1506, xz, probe_tile=256)
        engine.phase2_accum(gb, 0, M, inv, plain_tile, xz, meansq, N)
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        v0 += plain_tile
        est = meansq[:, 0] / float(v0) - M / float(N - 1)
        e = float(torch.mean(torch.abs(est - exact_d) / torch.abs(exact_d)).item())
        if v0 % (4 * plain_tile) == 0:
            trace.append({"probes": v0, "seconds": round(elapsed, 3), "err": round(e, 5)})
        if e <= 0.01:
            reached = {"probes": v0, "seconds": round(elapsed, 4), "err": round(e, 5)}
            break
    gb.close()
    out["plain_to_1pct"] = reached
    out["plain_trace"] = trace
    if reached is not None and out["errors"]["cv_b8"] <= 0.01:
        out["speedup_to_1pct"] = round(reached["seconds"] / t_cv, 1)
    log(f"plain to 1%: {reached}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=100_000)
    ap.add_argument("--m", type=int, default=500_000)
    ap.add_argument("--b", type=int, default=64)
    ap.add_argument("--limbs", type=int, default=5)
    ap.add_argument("--impute", default="mean")
    ap.add_argument("--cpu-snps", type=int, default=20000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-exact", action="store_true")
    ap.add_argument("--no-cv", action="store_true", help="skip the learned control-variate measurement")
    ap.add_argument("--no-tune", action="store_true", help="keep the built-in schedule heuristics")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay timing of the same steps")
    ap.add_argument("--only", choices=["learned_cv", "partitioned", "gxe", "windowed", "wall_to_1pct",
                                       "cpu_mailman_c1", "h2_end_to_end", "dropin_loop", "gxe_direct_loop",
                                       "vendor_int8_gemm"],
                    help="run only one extra measurement and print it")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the 1%%-error wall-time run and the host->HBM ingest measurement")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a plain 1 GiB read + 1 GiB write elementwise kernel three times (the reference "
                         "point that scales FETCH_SIZE / WRITE_SIZE counter readings to bytes)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU (the multi-GPU design is documented, not built); "
              "run with --gpus 1", file=sys.stderr)
        return 2

    import torch

    from summit_amd import engine, ldscore

    dev = engine.default_device()
    if args.copy_calib:
        x = torch.ones(1 << 28, dtype=torch.int32, device=dev)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        log("copy calibration: 3 x (1 GiB read + 1 GiB write)")
        return 0
    if args.only == "learned_cv":
        print(json.dumps({"learned_cv": learned_cv(engine, ldscore, torch, dev)}), flush=True)
        return 0
    if args.only == "partitioned":
        print(json.dumps({"partitioned": partitioned(engine, ldscore, torch)}), flush=True)
        return 0
    if args.only == "dropin_loop":
        print(json.dumps({"dropin_loop": dropin_loop(engine, ldscore, torch)}), flush=True)
        return 0
    if args.only == "gxe_direct_loop":
        print(json.dumps({"gxe_direct_loop": gxe_direct_loop(engine, torch)}), flush=True)
        return 0
    if args.only == "vendor_int8_gemm":
        print(json.dumps({"vendor_int8_gemm": vendor_int8_gemm(torch, dev)}), flush=True)
        return 0
    if args.only == "h2_end_to_end":
        print(json.dumps({"h2_end_to_end": h2_end_to_end(engine, ldscore, torch)}), flush=True)
        return 0
    if args.only == "gxe":
        print(json.dumps({"gxe": gxe_rate(engine, torch)}), flush=True)
        return 0
    if args.only == "windowed":
        print(json.dumps({"windowed": windowed_rate(engine, torch)}), flush=True)
        return 0
    if args.only == "cpu_mailman_c1":
        print(json.dumps({"cpu_mailman_c1": cpu_mailman_c1(engine, torch)}), flush=True)
        return 0
    if args.only == "wall_to_1pct":
        print(json.dumps({"wall_to_1pct": wall_to_one_percent(engine, ldscore, torch, dev)}), flush=True)
        return 0
    N, M, B = args.n, args.m, args.b
    seed = 1303
    t_gen = time.perf_counter()
    geno = engine.Genotype.synthetic(N, M, seed=303, maf_lo=0.01, maf_hi=0.5, miss_rate=0.005, block_len=100,
                                     copy_prob=0.9)
    geno.prepare(ddof=1, impute=args.impute, impute_seed=ldscore.impute_seed_from_root(seed))
    torch.cuda.synchronize()
    log(f"genotypes N={N} M={M} generated+prepared in {time.perf_counter() - t_gen:.1f} s, "
        f"missing calls {geno.n_missing()}")

    # schedule: a cached choice for this device and shape, else tune now (outside the timed region)
    from summit_amd import autotune
    tuned = None
    if not args.no_tune:
        tuned = {"phase1": autotune.apply_cached(N, M, 1, B, args.limbs),
                 "phase2": autotune.apply_cached_phase2(M, N, 1, B, args.limbs)}
        if tuned["phase1"] is None:
            tuned["phase1"] = autotune.tune_phase1(geno, B, 1, args.limbs, seed=seed)
        if tuned["phase2"] is None:
            tuned["phase2"] = autotune.tune_phase2(geno, B, 1, args.limbs, seed=seed)
        log(f"autotune: {tuned}")

    annot = torch.ones((M, 1), dtype=torch.float64, device=dev)
    inv = engine.resvar_inv(geno, 0, M)
    meansq = torch.zeros((M, 1), dtype=torch.float64, device=dev)
    xz = torch.empty((B, N), dtype=torch.float64, device=dev)

    def step(i: int) -> None:
        xz.zero_()
        engine.phase1_xz(geno, 0, M, 1000, annot, inv, i * B, B, seed, xz, limbs=args.limbs)
        engine.phase2_accum(geno, 0, M, inv, B, xz, meansq, N, limbs=args.limbs)

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    log(f"warmup {args.warmup} steps done")
    engine.profile_enable(True)
    engine.profile_reset()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(args.warmup + i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    p1_ms, p1_n = engine.profile_read("p1_contract")
    p2_ms, p2_n = engine.profile_read("p2_contract")
    fin_ms, fin_n = engine.profile_read("p1_finalize")
    pm_ms, pm_n = engine.profile_read("p2_miss")
    engine.profile_enable(False)
    ms_per_step = (t1 - t0) * 1000.0 / args.steps
    pairs = 2.0 * N * M * B
    value = pairs / (ms_per_step / 1000.0) / 1e9
    log(f"{args.steps} steps: {ms_per_step:.2f} ms/step, p1 contract {p1_ms / max(p1_n, 1):.2f} ms, "
        f"p2 contract {p2_ms / max(p2_n, 1):.2f} ms")

    # the same steps replayed from one HIP graph (each replay = the next probe tile)
    graph = None
    if not args.no_graph:
        gms = torch.zeros((M, 1), dtype=torch.float64, device=dev)
        gt = ldscore.GraphedTiles(geno, B, seed, annot, inv, gms, limbs=args.limbs,
                                  v_start=(args.warmup + args.steps) * B)
        gt.eager()
        gt.capture()
        gt.replay()
        torch.cuda.synchronize()
        tg0 = time.perf_counter()
        for _ in range(args.steps):
            gt.replay()
        torch.cuda.synchronize()
        g_ms = (time.perf_counter() - tg0) * 1000.0 / args.steps
        graph = {"ms_per_step": round(g_ms, 3), "gop_s": round(2.0 * N * M * B / (g_ms / 1000.0) / 1e9, 1),
                 "steps": args.steps}
        log(f"graph replays: {graph}")
        del gt, gms

    # dominant kernel against its roofline: algorithmic int ops (2 per genotype-probe MAC) per launch
    if p2_ms >= p1_ms:
        kname, kms, kn = "k_contract phase 2 (X^T(X.Z), fused Work^2)", p2_ms, p2_n
    else:
        kname, kms, kn = "k_contract phase 1 (X.Z)", p1_ms, p1_n
    avg_ms = kms / max(kn, 1)
    ops = 2.0 * N * M * B
    achieved = ops / (avg_ms / 1000.0) / 1e12
    # HBM bytes per launch of the dominant kernel from the counter runs of
    # profiles/traffic.json, reported only when they were read from this build
    traffic = None
    traffic_note = "no counter run recorded for this shape"
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        entry = tj.get("kernels", {}).get(f"N={N},M={M},B={B},limbs={args.limbs},impute={args.impute}")
        if entry is None:
            pass
        elif tj.get("source_sha256") != kernel_source_sha():
            traffic_note = "counter run of another build of the kernels: not reported"
        else:
            key = "p2_contract" if kname.startswith("k_contract phase 2") else "p1_contract"
            traffic = entry.get(key)
            traffic_note = tj.get("measured", "")
    # int8 MFMA multiply-adds the kernel issues per algorithmic genotype-probe pair: one per limb of
    # the real-valued operand, doubled in phase 2 when the missing-call plane is contracted too (with
    # sparse missing calls their sums are gathered over per-SNP lists instead: p2_miss_sums)
    miss_plane = args.impute == "mean" and geno.n_missing() > 0 and pm_n == 0
    per_pair = args.limbs * (2 if (kname.startswith("k_contract phase 2") and miss_plane) else 1)
    roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": INT8_PEAK_TOPS, "unit": "TFLOP/s",
                "frac": round(achieved / INT8_PEAK_TOPS, 4), "traffic": traffic, "traffic_source": traffic_note,
                "kernel": kname,
                "int8_macs_per_pair": per_pair, "mfma_pipe_frac": round(achieved * per_pair / INT8_PEAK_TOPS, 4),
                "avg_ms": round(avg_ms, 3),
                "phase_ms": {"p1_contract": round(p1_ms / max(p1_n, 1), 3),
                             "p1_missing_finalize": round(fin_ms / args.steps, 3),
                             "p2_contract": round(p2_ms / max(p2_n, 1), 3),
                             "p2_miss_sums": round(pm_ms / args.steps, 3)}}

    extra = {}
    if not args.no_exact:
        # exact integer path on the same genotypes: U = G.Z (int8 +-1), W = G^T U (3 limbs), bit-exact
        g = torch.Generator(device="cpu").manual_seed(5)
        z = (torch.randint(0, 2, (M, B), generator=g, dtype=torch.int8) * 2 - 1).to(dev)
        u = engine.gz_i32(geno, z)
        w = engine.gtu_i64(geno, u)
        torch.cuda.synchronize()
        engine.profile_enable(True)
        engine.profile_reset()
        for _ in range(3):
            u = engine.gz_i32(geno, z)
            w = engine.gtu_i64(geno, u)
        torch.cuda.synchronize()
        gz_ms, gz_n = engine.profile_read("gz")
        gtu_ms, gtu_n = engine.profile_read("gtu")
        engine.profile_enable(False)
        gz_avg, gtu_avg = gz_ms / gz_n, gtu_ms / gtu_n
        bed_bytes = M * ((N + 3) // 4)
        extra["exact_int"] = {
            "gz_ms": round(gz_avg, 3), "gtu_ms": round(gtu_avg, 3),
            "gop_s": round(pairs / ((gz_avg + gtu_avg) / 1000.0) / 1e9, 1),
            "gz_hbm_gbs": round(bed_bytes / (gz_avg / 1000.0) / 1e9, 1),
            "gtu_tops": round(2.0 * N * M * B / (gtu_avg / 1000.0) / 1e12, 1)}
        del u, w, z

    if not args.no_extras:
        # the same step with the reference's float32 Gaussian probes (host std::normal_distribution<float>
        # per SNP block, real coefficients on the same int8-limb contraction)
        def gstep(i: int) -> None:
            xz.zero_()
            engine.phase1_xz(geno, 0, M, 1000, annot, inv, i * B, B, seed, xz, rand_dist="gaussian", probe_fp32=True,
                             limbs=args.limbs)
            engine.phase2_accum(geno, 0, M, inv, B, xz, meansq, N, limbs=args.limbs)
        gstep(0)
        torch.cuda.synchronize()
        tg = time.perf_counter()
        for i in range(3):
            gstep(1 + i)
        torch.cuda.synchronize()
        gms = (time.perf_counter() - tg) * 1000.0 / 3
        extra["gaussian_f32"] = {"ms_per_step": round(gms, 3), "gop_s": round(pairs / (gms / 1000.0) / 1e9, 1),
                                 "workload": f"N={N} M={M} B={B} float32 Gaussian probes, 3 steps"}
        log(f"gaussian f32: {extra['gaussian_f32']}")
        extra["ingest"] = ingest_rate(engine, torch, geno, min(M, 100_000))
        log(f"ingest: {extra['ingest']}")
        if not args.no_cpu:
            extra["cpu_mailman_c1"] = cpu_mailman_c1(engine, torch)
            log(f"cpu mailman (C1 shape): {extra['cpu_mailman_c1']}")

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(geno, min(args.cpu_snps, M), B, seed)
        except Exception as e:  # the baseline is reported, never the metric
            log(f"cpu baseline failed: {e}")
    geno.close()
    if not args.no_extras:
        extra["wall_to_1pct"] = wall_to_one_percent(engine, ldscore, torch, dev)
        extra["windowed"] = windowed_rate(engine, torch)
        log(f"windowed: {extra['windowed']}")
        extra["gxe"] = gxe_rate(engine, torch)
        log(f"gxe: {extra['gxe']}")
    if not args.no_extras:
        extra["partitioned"] = partitioned(engine, ldscore, torch)
        log(f"partitioned: {extra['partitioned']}")
        extra["dropin_loop"] = dropin_loop(engine, ldscore, torch)
        log(f"drop-in loop: {extra['dropin_loop']}")
        extra["gxe_direct_loop"] = gxe_direct_loop(engine, torch)
        log(f"GxE drop-in loop: {extra['gxe_direct_loop']}")
        extra["h2_end_to_end"] = h2_end_to_end(engine, ldscore, torch)
        log(f"h2 end to end: {extra['h2_end_to_end']}")
        vg = vendor_int8_gemm(torch, dev)
        # the contraction's own int8 MFMA rate (limb products issued per second) beside the library's
        vg["phase2_contract_tops"] = round(roofline["achieved"] * roofline["int8_macs_per_pair"], 1)
        extra["vendor_int8_gemm"] = vg
        log(f"vendor int8 gemm: {vg}")
    if not args.no_extras and not args.no_cv:
        extra["learned_cv"] = learned_cv(engine, ldscore, torch, dev)
    line = {
        "metric": "genotype\u00b7probe Gop/s at N=100k,M=500k,B=64; wall time to 1% LD-score rel err",
        "value": round(value, 1),
        "unit": "Gop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int8-limb MFMA (i32 acc) + f64",
        "data": "synthetic (GPU block-Markov generator, seed 303)",
        "config": {"workload": f"genome-wide stochastic LD scores, N={N} M={M} B={B} Rademacher, "
                               f"block-Markov genotypes, 0.5% missing, {args.impute} imputation, "
                               f"step_size 1000, {args.limbs} int8 limbs",
                   "n": N, "m": M, "b": B, "limbs": args.limbs, "impute": args.impute},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "autotune": tuned,
        "graph_replay": graph,
        **extra,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
