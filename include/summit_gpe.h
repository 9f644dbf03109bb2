/*
 * summit_gpe.h — C ABI of libsummit_gpe.so, the MI355X (gfx950) genotype-probe
 * engine behind SUMMIT's stochastic LD scores.
 *
 * Conventions
 *   - Every entry point returns 0 on success and a nonzero status otherwise;
 *     gpe_last_error() returns the message of the calling thread's last
 *     failure (the reference raises std::runtime_error with the same texts,
 *     surfaced as Python RuntimeError by its bindings).
 *   - Pointers named *_dev are device (HBM) pointers; *_host are host
 *     pointers.  Sizes are element counts.  Matrices are described by a
 *     leading dimension; "column-major (N, Q)" means element (n, q) at
 *     q * ld + n, matching the reference's Fortran arrays.
 *   - stream is a hipStream_t (NULL = legacy default stream).  Calls enqueue
 *     work and return; results are complete after the stream is synchronized,
 *     except where a host output is written (those calls synchronize).
 *   - A gpe_geno handle owns one genotype matrix resident in HBM in two packed
 *     2-bit layouts (SNP-major and sample-major) plus its per-SNP statistics.
 *
 * Reference paths below are relative to the SUMMIT repository root.
 */
#ifndef SUMMIT_GPE_H
#define SUMMIT_GPE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gpe_geno gpe_geno;

enum {
    GPE_OK = 0,
    GPE_EINVAL = 1,   /* bad argument / shape (reference: std::runtime_error) */
    GPE_EDEVICE = 2,  /* HIP runtime failure */
    GPE_EINTERNAL = 3
};

enum { GPE_IMPUTE_MEAN = 0, GPE_IMPUTE_HWE = 1 };      /* src/native/common/genotype.hpp:33-36 */
/* Probe distributions (src/native/gwldcore.cpp:1681-1716): the reference draws
 * in the precision of the caller's arrays, so Gaussian / spherical streams come
 * in a double and a float version (float variates, stored as double). */
enum {
    GPE_RAND_RADEMACHER = 0,
    GPE_RAND_GAUSSIAN = 1,
    GPE_RAND_SPHERICAL = 2,
    GPE_RAND_GAUSSIAN_F32 = 3,
    GPE_RAND_SPHERICAL_F32 = 4
};
enum { GPE_LAYOUT_SNP_RAW = 0, GPE_LAYOUT_SNP_WORK = 1, GPE_LAYOUT_SAMPLE_WORK = 2 };

/* Message of the calling thread's last failure ("" if none). */
const char* gpe_last_error(void);

/* Library version string. */
const char* gpe_version(void);

/* Select the HIP device used by subsequent calls of this thread. */
int gpe_set_device(int device);

/* ------------------------------------------------------------------------
 * Genotype residency
 * ------------------------------------------------------------------------ */

/* Allocate a zeroed device genotype matrix of n_samples x n_snps.
 * Replaces the mmap'd BED cache entry (src/native/common/genotype.cpp:350-410)
 * with an HBM-resident matrix. */
int gpe_geno_create(int64_t n_samples, int64_t n_snps, gpe_geno** out);

/* Copy SNP rows [snp_begin, snp_begin + snp_count) of a host BED body (the
 * bytes after the 3-byte magic; host_stride = ceil(n_total/4) bytes per SNP)
 * into the handle, keeping only the sorted global sample ids rows_host
 * (NULL = all n_total samples; then n_total must equal n_samples).
 * Follows the decode of src/native/common/genotype.cpp:474-574. */
int gpe_geno_upload_bed(gpe_geno* g, const uint8_t* bed_host, int64_t n_total, int64_t host_stride,
                        int64_t snp_begin, int64_t snp_count, const int64_t* rows_host, void* stream);

/* Fill the handle with synthetic genotypes (no reference counterpart:
 * workload generator).  copy_prob = 0 gives i.i.d. Binomial(2, MAF). */
int gpe_geno_synthesize(gpe_geno* g, uint64_t seed, double maf_lo, double maf_hi, double miss_rate,
                        int32_t block_len, double copy_prob, void* stream);

/* Derive the working codes and per-SNP statistics: observed counts, mean and
 * inverse sd (src/native/common/genotype.cpp:629-689), HWE hash completion of
 * missing calls for impute_mode == GPE_IMPUTE_HWE
 * (src/native/common/genotype.cpp:122-184, :576-595; snp_offset = global
 * index of the handle's SNP 0, global_rows_host = global sample ids or NULL),
 * the sample-major layout, and sorted per-sample missing-call lists. */
int gpe_geno_prepare(gpe_geno* g, int32_t ddof, int32_t impute_mode, uint64_t impute_seed,
                     int64_t snp_offset, const int64_t* global_rows_host, void* stream);

/* Host copies of the per-SNP statistics (any pointer may be NULL).  Counts
 * are of the working codes (after HWE completion when prepared that way). */
int gpe_geno_stats(gpe_geno* g, int32_t* nobs_host, int64_t* sum_host, int64_t* sumsq_host,
                   double* mean_host, double* inv_sd_host);

/* Download one packed layout (rows of `host_stride` bytes; SNP layouts have
 * n_snps rows of ceil(n_samples/4) used bytes, the sample layout n_samples rows
 * of ceil(n_snps/4) used bytes). */
int gpe_geno_download(gpe_geno* g, int32_t layout, uint8_t* host, int64_t host_stride);

/* Download rows [row_begin, row_begin + row_count) of one packed layout. */
int gpe_geno_download_rows(gpe_geno* g, int32_t layout, int64_t row_begin, int64_t row_count, uint8_t* host,
                           int64_t host_stride);

/* Sizes of the handle. */
int gpe_geno_shape(gpe_geno* g, int64_t* n_samples, int64_t* n_snps, int64_t* n_missing);

int gpe_geno_free(gpe_geno* g);

/* ------------------------------------------------------------------------
 * Exact integer products over the working genotype counts g (missing -> 0)
 * ------------------------------------------------------------------------ */

/* U[b][n] (+)= sum_{i in [snp_begin, snp_end)} g[i][n] * Z[i - snp_begin][b]
 * Z: int8 row-major (snp_end - snp_begin) x n_cols (ldz >= n_cols), device.
 * U: int32 column-major (n_samples, n_cols), ldu >= n_samples, device.
 * |U| must stay below 2^31 (checked from max|Z| and the SNP count).
 * The unit-weight form of phase 1 (src/native/gwldcore.cpp:1809-1834). */
int gpe_gz_i32(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int8_t* z_dev, int64_t ldz,
               int32_t n_cols, int32_t* u_dev, int64_t ldu, int32_t accumulate, void* stream);

/* W[i - snp_begin][b] = sum_n g[i][n] * U[b][n], exact int64, row-major
 * (ldw >= n_cols).  U as produced by gpe_gz_i32 with |U| < 2^23 (three int8
 * limbs).  The unit-weight form of phase 2's GEMM
 * (src/native/gwldcore.cpp:2485-2491). */
int gpe_gtu_i64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int32_t* u_dev, int64_t ldu,
                int32_t n_cols, int64_t* w_dev, int64_t ldw, void* stream);

/* ------------------------------------------------------------------------
 * Stochastic LD-score phases (reference estimator, FP64 semantics)
 * ------------------------------------------------------------------------ */

/* inv_out_dev[j - snp_begin] = 1/sqrt(max(||x_j||^2/(N-1), eps)) for the
 * standardized SNPs of the prepared handle, no covariates.
 * Replaces compute_inv_resvar_from_geno_block (src/native/gwldcore.cpp:3220-3333)
 * as used by precompute_residual_variances_bed (:3335-3566) and the fused
 * inv_out of phase 1. */
int gpe_resvar_inv(gpe_geno* g, int64_t snp_begin, int64_t snp_end, double eps, double* inv_out_dev,
                   void* stream);

/* Phase 1, X.Z: for SNPs [snp_begin, snp_end) in blocks of step_size (block b
 * draws its probes from make_seed(root_seed, block_start, v_start)),
 *   Xz[:, k*v_count + v] (+)= sum_i x_i * inv_i * sqrt(a_ik) * z_iv
 * annot_dev: (snp_end - snp_begin) x n_bins row-major FP64; inv_dev: FP64
 * per SNP; Xz: column-major (n_samples, n_bins * v_count), ldxz >= n_samples.
 * bin_mode_dev (may be NULL = accumulate every bin): per bin 1 = overwrite,
 * 2 = accumulate.  limbs: 5 or 6 int8 limbs per coefficient.
 * Replaces phase1_compute_Xz_bed_chunk (src/native/gwldcore.cpp:1582-1923,
 * exported at :3673-3688); one call may span many reference blocks. */
int gpe_phase1_xz(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size,
                  const double* annot_dev, int32_t n_bins, const double* inv_dev, int32_t v_start,
                  int32_t v_count, int32_t rand_dist, uint64_t root_seed, double* xz_dev, int64_t ldxz,
                  const uint8_t* bin_mode_dev, int32_t limbs, void* stream);

/* gpe_phase1_xz over several of the reference's V-tiles at once: probe
 * columns [t*probe_tile, (t+1)*probe_tile) carry the streams of V-tile
 * v_start + t*probe_tile (make_seed(root, block start, v_start + t*probe_tile)),
 * exactly as separate calls with v_count = probe_tile would, in one traversal.
 * The reference's probe values depend on its V-tile split
 * (src/summit/ldscore/gw_ldscore.py:687-697); this lets a wide traversal keep a
 * narrow split, and gives the GPU one probe stream per (block, tile) to run in
 * parallel. */
int gpe_phase1_xz_tiled(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size, const double* annot_dev,
                        int32_t n_bins, const double* inv_dev, int32_t v_start, int32_t v_count, int32_t probe_tile,
                        int32_t rand_dist, uint64_t root_seed, double* xz_dev, int64_t ldxz,
                        const uint8_t* bin_mode_dev, int32_t limbs, void* stream);

/* gpe_phase1_xz_tiled with Rademacher probes whose first probe id (the
 * v_start of the reference's stream seed) is read on the device from
 * *v_start_dev when the kernels run.  With gpe_add_i32 advancing it, a probe
 * tile (phase 1, phase 2, advance) can be captured once in a HIP graph and
 * replayed tile after tile.  Under capture the workspaces and the device copy
 * of the SNP-block starts must already exist: run the same call once
 * outside capture first (otherwise the call fails with that message). */
int gpe_phase1_xz_indirect(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size,
                           const double* annot_dev, int32_t n_bins, const double* inv_dev,
                           const int32_t* v_start_dev, int32_t v_count, int32_t probe_tile, uint64_t root_seed,
                           double* xz_dev, int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream);

/* *value_dev += increment on the device (stream-ordered; graph-capturable). */
int gpe_add_i32(int32_t* value_dev, int32_t increment, void* stream);

/* Phase 2, X^T(X.Z): for SNPs j in [snp_begin, snp_end) and bins k,
 *   meansq[j][k] += sum_v (x_j . Xz[:, k*tile_nvecs + v])^2 * inv_j^2/(N_denom-1)^2
 * with the non-finite / extreme-update guards of src/native/gwldcore.cpp:295-417
 * (extreme_warn <= 0 disables the extreme test), and optional probe fourth-moment
 * sums (mc_aggregate_sumsq_dev[k], mc_per_snp_sumsq_dev[j][k]; NULL to skip).
 * meansq: row-major (n_snps_total, n_bins) with ld_meansq >= n_bins, indexed by
 * absolute SNP j.  inv_left_dev: per SNP of the range.  guard_counts_host
 * (may be NULL) receives {nonfinite, extreme} counts (the call synchronizes).
 * A NaN or infinity anywhere in a column of Xz makes that column's product with
 * every SNP non-finite, as in the reference's dense GEMM, so the updates of its
 * bin are counted as non-finite and dropped.
 * Replaces phase2_accum_XtXz_bed (src/native/gwldcore.cpp:2329-2649, exported
 * at :3759-3770) and, with HWE-prepared handles, the Mailman variants
 * (:2651-3119). */
int gpe_phase2_accum(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* inv_left_dev,
                     int32_t tile_nvecs, const double* xz_dev, int64_t ldxz, int32_t n_bins,
                     int32_t n_denom, double* meansq_dev, int64_t ld_meansq, double* mc_aggregate_sumsq_dev,
                     double* mc_per_snp_sumsq_dev, double extreme_warn, int32_t drop_extreme, int32_t limbs,
                     unsigned long long* guard_counts_host, void* stream);

/* gpe_phase2_accum with the centring taken from the caller's column sums:
 * Work_jc = inv_sd_j (g_j . Xz_c - mu_j sum_xz_dev[c]) (minus the missing-call
 * terms of a mean-imputed handle) instead of the exact sums of Xz, for
 * c < n_bins * tile_nvecs (device FP64).  Replaces
 * phase2_accum_XtXz_bed_mailman_rowmajor (src/native/gwldcore.cpp:2919-3119,
 * exported at :3800-3809), whose per-column centring uses the caller's sum_Xz
 * (:2965-2967, :874-880); its callers pass the column sums of the projected
 * X.Z (src/summit/ldscore/gw_ldscore.py:2010-2016).  A non-finite sum makes its
 * column's updates non-finite (counted and dropped like any other). */
int gpe_phase2_accum_colsums(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* inv_left_dev,
                             int32_t tile_nvecs, const double* xz_dev, int64_t ldxz, int32_t n_bins,
                             const double* sum_xz_dev, int32_t n_denom, double* meansq_dev, int64_t ld_meansq,
                             double* mc_aggregate_sumsq_dev, double* mc_per_snp_sumsq_dev, double extreme_warn,
                             int32_t drop_extreme, int32_t limbs, unsigned long long* guard_counts_host,
                             void* stream);

/* Dense products with the standardized genotype matrix X (N x M, the handle's
 * working codes, x_in = (g_in - mu_i) inv_sd_i, 0 where missing in mean mode):
 *
 *   gpe_xc_f64:  out[q][n] (+)= sum_{i in [snp_begin, snp_end)} x_in * c[i - snp_begin][q]
 *                c row-major (L, n_cols) FP64; out column-major (N, n_cols) with ldo >= N.
 *   gpe_xtv_f64: out[j - snp_begin][q] = sum_n x_jn * v[q][n]
 *                v column-major (N, n_cols) (ldv >= N); out row-major (L, n_cols).
 *
 * The two GEMMs of apply_grm_bed_panel (src/native/gwldcore.cpp:1227-1252:
 * coef = X_blk^T V, out += X_blk (diag(inv^2/M) coef)) and of the GxE
 * source / target products; limbs 5 or 6 as for the phases. */
int gpe_xc_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* c_dev, int64_t ldc, int32_t n_cols,
               double* out_dev, int64_t ldo, int32_t accumulate, int32_t limbs, void* stream);
int gpe_xtv_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* v_dev, int64_t ldv, int32_t n_cols,
                double* out_dev, int64_t ldo, int32_t limbs, void* stream);

/* Exact windowed LD scores for SNPs [snp_begin, snp_end) of one chromosome:
 *   out[j][b] = sum_{i : |bp_i - bp_j| <= window_bp} r~2_ij a[i][b],
 *   r~2 = r^2 - (1 - r^2)/(N - 2),  r the correlation of the ddof-0
 * standardized, N-restandardized SNPs (exact integer cross products on int8
 * MFMA).  window_snps > 0 replaces the BP window by |i - j| <= window_snps
 * (bp_host may then be NULL).  annot_dev, out_dev: row-major (m, n_bins), m =
 * snp_end - snp_begin.  Replaces compute_windowed_ld_chr without covariates
 * (src/native/winldcore.cpp:783-931, r^2 and masks :409-594). Synchronizes. */
int gpe_windowed_ld(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                    int64_t window_snps, const double* annot_dev, int32_t n_bins, double* out_dev, void* stream);

/* gpe_windowed_ld summed over one side of each SNP's window: side 1 takes the
 * window SNPs before j, side 2 those after j (j itself in neither; side 0 is
 * gpe_windowed_ld).  No counterpart in the reference: the asymmetric in-window
 * features of the learned LD-score predictor (how LD decays on each side). */
int gpe_windowed_ld_side(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                         int64_t window_snps, int32_t side, const double* annot_dev, int32_t n_bins, double* out_dev,
                         void* stream);

/* gpe_windowed_ld with covariates, as compute_windowed_ld_chr with C and R
 * (src/native/winldcore.cpp:783-931, panels :341-372): each SNP's ddof-0
 * standardized column x is projected, y = x - C (R x), and restandardized over
 * all N rows (centred, population variance 1), r_ij the correlation of the
 * projected columns, r~2 = r^2 - (1 - r^2)/(N - p - 1).  C column-major (N, p)
 * (c_dev[k * ldc + n]), R row-major (p, N) (r_dev[k * ldr + n]), device FP64.
 * The exact integer Gram gets a rank-2p FP64 correction per pair.  Synchronizes. */
int gpe_windowed_ld_cov(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                        int64_t window_snps, const double* annot_dev, int32_t n_bins, const double* c_dev,
                        int64_t ldc, const double* r_dev, int64_t ldr, int32_t p, double* out_dev, void* stream);

/* The same for n_chr chromosomes in ONE banded-Gram launch (every
 * chromosome's tile pairs in one work list): chromosome c is SNPs
 * [chr_bounds_host[2c], chr_bounds_host[2c+1]) (sorted, non-overlapping);
 * rows of bp_host, annot_dev and out_dev are SNP ids minus chr_bounds_host[0]
 * (rows between chromosomes are left 0).  The reference runs
 * compute_windowed_ld_chr once per chromosome (src/summit/ldscore/win_ldscore.py:974);
 * results equal those per-chromosome calls.  Synchronizes. */
int gpe_windowed_ld_multi(gpe_geno* g, int32_t n_chr, const int64_t* chr_bounds_host, const int64_t* bp_host,
                          double window_bp, int64_t window_snps, const double* annot_dev, int32_t n_bins,
                          double* out_dev, void* stream);

/* GxE probe streams (src/summit/ldscore/gwe_ldscore.py:2484-2505, native
 * restatement src/native/gxeldcore.cpp:10002-10138): block b (SNPs
 * [blk_starts[b], next start or blk_end)) and probe v draw NumPy Philox4x64-10
 * with key keys_host[(b*v_count + v)*2 .. +1] (NumPy's seeding of
 * Philox(_make_seed(root, block start, probe id))); value = +-1 from the high
 * bit of each generated byte.  z_dev: int8 row-major, row 0 = blk_starts[0]. */
int gpe_philox_rademacher(const uint64_t* keys_host, const int64_t* blk_starts_host, int32_t n_blocks, int64_t blk_end,
                          int32_t v_count, int8_t* z_dev, int64_t ldz, void* stream);

/* Phase 1 with caller-provided int8 probes z_dev (row-major (L, v_count), row 0
 * = snp_begin), otherwise as gpe_phase1_xz.  Probe values must be -1, 0 or +1
 * (Rademacher signs): the missing-call terms keep each probe as a 2-bit sign and
 * all columns of a bin share one fixed-point scale.  Outside graph capture the
 * call checks this on the device and returns GPE_EINVAL otherwise (the call then
 * synchronizes); under capture the caller guarantees it.  The GxE sources
 * (src/native/gxeldcore.cpp:7418-7576) use it with their Philox probes. */
int gpe_phase1_xz_z8(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* annot_dev, int32_t n_bins,
                     const double* inv_dev, const int8_t* z_dev, int64_t ldz, int32_t v_count, double* xz_dev,
                     int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream);

/* Phase 1 with caller-provided real-valued probes z_dev (FP64 row-major
 * (L, v_count), row 0 = snp_begin), otherwise as gpe_phase1_xz: the GxE
 * sources with Gaussian or spherical NumPy Philox probes
 * (src/summit/ldscore/gwe_ldscore.py:2484-2505, src/native/gxeldcore.cpp:7418-7576). */
int gpe_phase1_xz_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* annot_dev, int32_t n_bins,
                      const double* inv_dev, const double* z_dev, int64_t ldz, int32_t v_count, double* xz_dev,
                      int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream);

/* Small-rank FP64 projection of n_cols columns of length n, in place:
 *   x_q <- x_q - A (B x_q),  then, when center != 0, x_q <- x_q - mean(x_q),
 * x column q at x_dev[q * ldx + i]; A n x p column-major (a_dev[i + k * lda]);
 * B p x n row-major (b_dev[k * ldb + i]).  With A = C, B = R it is the
 * covariate projection (I - C R) of src/native/gwldcore.cpp:1119-1132 and
 * :3134-3207; with A = Q, B = Q^T and centring the GxE projection off the
 * intercept and an orthonormal covariate basis (src/native/gxeldcore.cpp:7578-7675).
 * Fixed-order reductions; a per-device workspace is kept for the process. */
int gpe_project_f64(double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* a_dev, int64_t lda,
                    const double* b_dev, int64_t ldb, int32_t p, int32_t center, void* stream);

/* gpe_project_f64 without any allocation: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments, kernels
 * and results as gpe_project_f64; its partial results live in one fixed 4 MiB
 * device array per GPU, held until the process ends, so a call needs
 * (max(p,1) + 2) * n_cols + max(p,1) <= 524288 (larger sizes return GPE_EINVAL).
 * Every call on one GPU shares that array: calls that overlap on one GPU (other
 * streams or threads) race and must be serialized by the caller. */
int gpe_project_f64_noalloc(double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* a_dev,
                            int64_t lda, const double* b_dev, int64_t ldb, int32_t p, int32_t center,
                            void* stream);

/* out_dev[q] = sum_{i < n} x_dev[q * ldx + i] for q < n_cols (FP64, fixed-order
 * reduction): the column sums of a row-major (N, Q) X.Z held column by column.
 * Replaces compute_col_sums_rowmajor (src/native/gwldcore.cpp:3121-3132,
 * exported at :3720-3721) and the sums of project_rowmajor_inplace_and_col_sums
 * (:3134-3171) after gpe_project_f64. */
int gpe_col_sums_f64(const double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, double* out_dev, void* stream);

/* t_dev[k * ldt + q] = sum_{i < n} b_dev[k * ldb + i] x_dev[q * ldx + i] for k < p, q < n_cols,
 * and, when sumsq_dev is not null, sumsq_dev[q] = sum_i x_dev[q * ldx + i]^2 (FP64,
 * fixed-order reductions): the basis coefficients Q^T S and squared norms of a
 * source panel whose ratio is the projected-source leakage diagnostic of
 * DirectContext (src/native/gxeldcore.cpp:8244-8287, used by
 * prepare_projected_sources :7578-7675 and target_block :7682-7812). */
int gpe_basis_coef_f64(const double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* b_dev, int64_t ldb,
                       int32_t p, double* t_dev, int64_t ldt, double* sumsq_dev, void* stream);

/* Per-SNP feature algebra of the GxE per-block backend (DirectContext,
 * src/native/gxeldcore.cpp:8033-8243), one thread per SNP, FP64.  Inputs for
 * n_snps SNPs: sums_dev rows k < 4 (sums_dev[k * lds + j]) the sums of squared
 * standardized genotypes weighted by 1, e, e^2, e^4 (gpe_weighted_sumsq);
 * moments_dev[j * ldm + c], c < 4q, the moments [Q, EQ, E^2 Q, E^3 Q]^T x_j
 * (gpe_xtv_f64) in the target standardization; r2_dev[j] the factor that takes
 * the sums to it; gram_dev / e2gram_dev the q x q matrices Q^T Q and Q^T E^2 Q
 * (element [b * q + a]); rank = N - q.  Output rows (out_dev[r * ldo + j]):
 * scale_x, scale_w, norm_x, norm_w, diag_nxe_x, diag_nxe_w, corr_xw, varx,
 * varw and the two projection-leakage ratios. */
int gpe_gxe_direct_features(const double* sums_dev, int64_t lds, const double* moments_dev, int64_t ldm,
                            int32_t q, const double* gram_dev, const double* e2gram_dev, const double* r2_dev,
                            double rank, int64_t n_snps, double* out_dev, int64_t ldo, void* stream);

/* out_dev[k * ldo + i] += sum_{snp_begin <= j < snp_end} xt_ij^2 w_dev[(j - snp_begin) + k * ldw]
 * for k < n_bins and every sample i (FP64), xt the standardized working genotype
 * (0 at a missing call under mean imputation) and, when p > 0, its projection
 * xt = x - C (R x) with C column-major (N, p) (ldc) and rx_dev[(j - snp_begin) *
 * ldrx + a] = (R x_j)_a (p <= 64).  With w = a_jk / resvar_j these are the
 * per-sample sums S of the fourth-moment (mu22) statistics of
 * precompute_residual_variances_bed(compute_mu22=True)
 * (src/native/gwldcore.cpp:3496-3521; mu22 = S^T S / (N Msum_a Msum_b), :3528-3544). */
int gpe_sq_rowsums_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* w_dev, int64_t ldw,
                       int32_t n_bins, const double* c_dev, int64_t ldc, const double* rx_dev, int64_t ldrx, int32_t p,
                       double* out_dev, int64_t ldo, void* stream);

/* out[e * ldo + (j - snp_begin)] = sum_n w[e * ldw + n] x_jn^2 for each of
 * n_weights weight vectors (standardized working codes): the ||e . x_j||^2
 * feature moments of the GxE interaction columns, every environment in one
 * pass over the genotypes per 16 weight vectors
 * (src/summit/ldscore/gwe_ldscore.py:2349-2353). */
int gpe_weighted_sumsq(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* w_dev, int32_t n_weights,
                       int64_t ldw, double* out_dev, int64_t ldo, void* stream);

/* Probe values of the reference's genome-wide streams for blocks starting at
 * blk_starts_host[0..n_blocks) (last block ends at blk_end): Rademacher on the
 * GPU; Gaussian and spherical with the host std::normal_distribution<T>.  z_host: row-major
 * (blk_end - blk_starts_host[0]) x v_count FP64.  Follows
 * src/native/gwldcore.cpp:445-458, :1679-1716. */
int gpe_probes(int32_t rand_dist, uint64_t root_seed, const int64_t* blk_starts_host, int32_t n_blocks,
               int64_t blk_end, int32_t v_start, int32_t v_count, double* z_host);

/* ------------------------------------------------------------------------
 * Learned LD-score predictor (no reference counterpart: the control variate
 * of the stochastic estimator).  A three-layer MLP in_dim -> 128 -> 128 -> 1
 * with GELU (erf form) after the first two layers; FP32 master weights used
 * as BF16, inputs and hidden activations as BF16, FP32 accumulation on bf16
 * MFMA, FP32 biases, Adam(W) with FP32 moments.
 *
 * Flat FP32 parameter vector (gpe_mlp_n_params elements): W1 (128 x 32,
 * row-major; columns >= in_dim are zero), b1 (128), W2 (128 x 128), b2 (128),
 * w3 (128), b3 (1).  Features are bf16 bit patterns (uint16): x_dev
 * row-major (rows, >= 32 columns, ldx a multiple of 8, zero beyond in_dim),
 * and for training also x^T (32 rows of ldxT >= rows).  Calls enqueue work on
 * `stream` and allocate nothing after gpe_mlp_create, so gpe_mlp_train_step
 * may be captured in a HIP graph.
 * ------------------------------------------------------------------------ */
typedef struct gpe_mlp gpe_mlp;

int gpe_mlp_n_params(int64_t* n);
/* Network on the current device with workspace for max_rows rows per step;
 * parameters start at zero (set them with gpe_mlp_set_params). */
int gpe_mlp_create(int32_t in_dim, int64_t max_rows, gpe_mlp** out);
int gpe_mlp_free(gpe_mlp* net);
/* Replace the parameters (synchronous); resets the Adam moments and step. */
int gpe_mlp_set_params(gpe_mlp* net, const float* params_host, int64_t n);
int gpe_mlp_get_params(gpe_mlp* net, float* params_host, int64_t n);
/* Gradient of the last gpe_mlp_train_step (before its update). */
int gpe_mlp_get_grad(gpe_mlp* net, float* grad_host, int64_t n);
/* Adam steps taken so far. */
int gpe_mlp_step_count(gpe_mlp* net, int32_t* step_host);
/* y_dev[r] = network(x_r) for r < rows (FP32); any row count (processed in
 * chunks of max_rows). */
int gpe_mlp_forward(gpe_mlp* net, const uint16_t* x_dev, int64_t ldx, int64_t rows, float* y_dev, void* stream);
/* One step on rows <= max_rows: loss = mean_r (y_r - t_r)^2 (written to
 * loss_dev[0] when non-NULL), its gradient, and when update != 0 one Adam(W)
 * update (decoupled weight decay) with bias correction by the step count. */
int gpe_mlp_train_step(gpe_mlp* net, const uint16_t* x_dev, int64_t ldx, const uint16_t* xT_dev, int64_t ldxT,
                       const float* target_dev, int64_t rows, float lr, float beta1, float beta2, float eps,
                       float weight_decay, int32_t update, float* loss_dev, void* stream);

/* ------------------------------------------------------------------------
 * Kernel timing (no reference counterpart: the reference's per-block phase
 * timers, src/native/gwldcore.cpp:225-258, print wall time to stderr; these
 * record HIP events around the contraction kernels on their own stream).
 * ------------------------------------------------------------------------ */
enum {
    GPE_PROF_P1_CONTRACT = 0, /* phase-1 packed contraction (X.Z) */
    GPE_PROF_P2_CONTRACT = 1, /* phase-2 packed contraction with fused Work^2 epilogue */
    GPE_PROF_GZ = 2,          /* exact G.Z */
    GPE_PROF_GTU = 3,         /* exact G^T U */
    GPE_PROF_GRAM = 4,        /* banded Gram of windowed LD */
    GPE_PROF_P1_FINALIZE = 5, /* phase-1 missing-call term: the records' finalize (all SNP-range launches of
                                 one traversal) or the missing-flag plane (quantization, traversal, combine) */
    GPE_PROF_P2_MISS = 6,     /* phase-2 missing-call sums gathered over the per-SNP missing lists */
    GPE_PROF_KINDS = 7
};
int gpe_profile_enable(int32_t on);
/* Accumulated milliseconds and launch count of one kernel kind (synchronizes
 * on the recorded events). */
int gpe_profile_read(int32_t kind, double* total_ms, int64_t* launches);
int gpe_profile_reset(void);

/* ------------------------------------------------------------------------
 * Schedule table (no reference counterpart: the reference's OpenMP/BLAS
 * blocking is fixed at build time).  Per shape key (rows of the contraction,
 * contraction length, padded columns, limbs) and knob, a value that overrides
 * the built-in heuristic; filled by the autotuner (summit_amd/autotune.py),
 * which times the candidates on the device.  value <= 0 removes the entry.
 *   GPE_SCHED_P1_SPLITS:        K splits of the phase-1 contraction (1..64)
 *   GPE_SCHED_P1_FINALIZE_KIB:  SNP-range size of the phase-1 missing-call
 *                               finalize, in KiB of records (64..65536; no
 *                               entry: 8 MiB for one bin, 64 MiB for several)
 *   GPE_SCHED_P1_CPB / P2_CPB:  128-value chunks staged per LDS slot and
 *                               barrier in the five-limb contractions (1 or 2;
 *                               2 when no entry)
 *   GPE_SCHED_P2_MISS:          phase-2 missing calls (mean imputation): 1 a
 *                               missing-flag MFMA plane, 2 exact sums gathered
 *                               over per-SNP missing lists (no entry: 2 up to
 *                               0.55% missing calls, else 1); results are identical
 *   GPE_SCHED_P2_MISS_PARTS:    sample ranges of the gathered sums (1, or a
 *                               multiple of 8 up to 64; 8 when no entry)
 *   GPE_SCHED_P1_MISS:          phase-1 missing calls with int8 probes: 1
 *                               per-call records, 2 a missing-flag MFMA plane
 *                               (no entry: a cost model of missing-call density
 *                               d and the call's SNP count L: the plane when
 *                               L (0.248 - 12.2 d) ps < 2.52 ns, e.g. for
 *                               L = 10,000 at 1%, records for L = 500,000 at 0.5%)
 * Phase-1 keys: rows = samples, length = SNPs of the call, columns = bins x
 * probes rounded up to 32 per bin.
 * ------------------------------------------------------------------------ */
enum {
    GPE_SCHED_P1_SPLITS = 0,
    GPE_SCHED_P1_FINALIZE_KIB = 1,
    GPE_SCHED_P1_CPB = 2,  /* chunks per barrier of the five-limb phase-1 contraction (1 or 2) */
    GPE_SCHED_P2_CPB = 3,  /* the same for phase 2; key rows = SNPs, length = samples */
    GPE_SCHED_P2_MISS = 4, /* phase-2 missing-call method (phase-2 key) */
    GPE_SCHED_P2_MISS_PARTS = 5, /* sample ranges of the gathered missing-call sums (phase-2 key) */
    GPE_SCHED_P1_MISS = 6, /* phase-1 missing-call method for int8 probes (phase-1 key) */
    GPE_SCHED_KNOBS = 7
};
int gpe_schedule_set(int32_t knob, int64_t rows, int64_t length, int32_t columns, int32_t limbs, int32_t value);
/* The value in force for a key (0 = the heuristic decides). */
int gpe_schedule_get(int32_t knob, int64_t rows, int64_t length, int32_t columns, int32_t limbs, int32_t* value);
int gpe_schedule_clear(void);
/* The built-in K-split rules the traversals use without a schedule entry (host
 * only, no device work): rule 0 the fill rule of the exact integer products (the
 * smallest split count whose last round of workgroups fills >= 90% of the
 * resident slots), rule 1 the phase-1 rule (splits until the grid holds 16 rounds
 * of workgroups while every split keeps >= 128 chunks, at most 16); both then
 * raised until no split exceeds max_chunks_per_split. */
int gpe_schedule_split_rule(int32_t rule, int64_t grid_xy, int64_t kcount, int64_t max_chunks_per_split,
                            int64_t slots, int32_t* splits);

#ifdef __cplusplus
}
#endif

#endif /* SUMMIT_GPE_H */
