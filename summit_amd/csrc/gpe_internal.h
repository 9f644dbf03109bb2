// Internal declarations shared by the kernel and API translation units of
// libsummit_gpe.so.  Not part of the public C ABI (include/summit_gpe.h).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace gpe {

// ---------------------------------------------------------------------------
// Device layout of a genotype matrix resident in HBM.
//
//   snp-major  : one row per SNP, 4 samples per byte (PLINK order inside a
//                byte: sample 4b+t in bits 2t..2t+1), rows padded to a
//                multiple of 64 bytes, row count padded to a multiple of 256.
//   sample-major: the transpose, one row per sample, 4 SNPs per byte, same
//                padding rules.  Both passes of the estimator contract along
//                the contiguous (packed) axis of one of these two layouts.
//
// Codes: 00 -> 0, 01 -> missing, 10 -> 1, 11 -> 2.  Padding codes are 00 and
// contribute nothing because every coefficient at a padded position is 0.
// ---------------------------------------------------------------------------
constexpr int kKChunk = 128;        // contraction-axis values per MFMA chunk
constexpr int kRowPad = 256;        // row-count padding of both layouts
constexpr int kXcds = 8;            // MI355X: accelerator complex dies, one L2 each

__host__ __device__ inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }
__host__ __device__ inline int64_t ceil_div(int64_t x, int64_t m) { return (x + m - 1) / m; }

struct Error : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define GPE_HIP(expr)                                                                     \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            throw ::gpe::Error(std::string("HIP error ") + hipGetErrorString(_e) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__) + " in " #expr);  \
    } while (0)

// Output modes of the packed-contraction kernel.
enum EpiMode : int {
    EPI_I32_ATOMIC = 0,  // NL == 1: atomically add int32 results into out_i32[c * ldo + r]
    EPI_I64_ATOMIC = 1,  // combine limbs, atomically add int64 into out_i64[c * ldo + r]
    EPI_I64_ROWMAJOR = 2,// combine limbs, store int64 to out_i64[r * ldo + c]
    EPI_P2 = 3,          // phase-2 epilogue: standardized Work, per-tile sums of Work^2, Work^4
    EPI_WORK = 4,        // phase-2 epilogue writing standardized Work to out_f64[r * ldo + c]
    EPI_I64_2P = 5,      // MISS: int64 row-major of the genotype plane to out_i64, missing plane to out2_i64
    EPI_I64_SQ = 6,      // squared-genotype plane (codes 00 -> 0, 01 -> 0, 10 -> 1, 11 -> 4), int64 row-major
    EPI_I64_ATOMIC_Q = 7,// missing-flag plane (01 -> 1), otherwise as EPI_I64_ATOMIC
    EPI_P2X = 8,         // EPI_P2 with the missing-call sums read from msum instead of a missing plane
};

struct ContractArgs {
    const uint8_t* packed;   // packed rows (either layout)
    int64_t stride;          // bytes per packed row
    int32_t tiled;           // 0: row-major rows; 1: chunk-tiled copy (tile_packed)
    int64_t r0, r1;          // rows [r0, r1) are produced
    const int8_t* coef;      // swizzled coefficient limbs
    int64_t coef_ctiles;     // c-tiles per K chunk in the swizzled layout
    int64_t kc0, kc1;        // K chunks [kc0, kc1)
    int64_t kc_per_split;    // K chunks per split (grid.z)
    int32_t ctiles_used;     // number of real c-tiles (per limb) in the output
    // outputs
    int32_t* out_i32;
    int64_t* out_i64;
    double* out_f64;
    int64_t ldo;
    // phase-2 epilogue inputs
    const double* mean;      // per row (SNP) mean
    const double* inv_sd;    // per row inv sd
    const double* col_scale; // per column 2^-shift
    const int64_t* col_sum;  // per column exact sum of quantized values
    double* s2;              // [row][ctile] sum of Work^2 over the tile's columns
    double* s4;              // [row][ctile] sum of Work^4 (may be null)
    int64_t ld_s;            // ctile count stride of s2/s4
    int has_miss;
    int32_t cols_out;        // I32/I64 epilogues: columns >= cols_out are not written
    int64_t out_row0;        // I32/I64 epilogues: output row index = r - out_row0
    int64_t* out2_i64;       // EPI_I64_2P: missing-plane output (same layout as out_i64)
    const int64_t* msum;     // EPI_P2X: exact missing-call sums, msum[(r - r0) * ld_ms + c]
    int64_t ld_ms;
};

// Banded Gram (windowed LD) arguments
struct GramArgs {
    const uint8_t* packed;
    int64_t stride;
    int64_t n;           // samples
    int64_t kc1;         // K chunks
    const int32_t* pairs;  // (ti, tj, chr) triples: 64-SNP tiles of chromosome chr, ti <= tj,
                           // counted from the chromosome's first SNP
    const int64_t* chr;    // (begin, end) absolute SNP bounds per chromosome
    int64_t row0;          // absolute SNP index of row 0 of bp, annot and out
    const double* mean;
    const double* scale;   // s_i
    const int64_t* sum;
    const int32_t* nobs;
    const int64_t* bp;     // relative to row0
    double window_bp;      // used when window_snps <= 0
    int64_t window_snps;
    const double* annot;   // row-major (rows, n_bins), relative to row0
    int n_bins;
    double* out;           // row-major (rows, n_bins), relative to row0
    double corr_dim;
    // covariates (cov_k > 0): r_ij = (s_i s_j dot + cov_a[i] . cov_b[j] - N ybar_i ybar_j) t_i t_j / N,
    // per-SNP rows of cov_k values indexed by absolute SNP minus row0
    const double* cov_a;
    const double* cov_b;
    int cov_k;
    const double* ybar;
    const double* tsc;
    // 0: out[j] sums over the whole window; 1: only window SNPs before j; 2: only after j
    int side;
};

// Launchers (gpe_kernels.hip)
// cpb: chunks per barrier (1, or 2 for the five-limb phase-1 / phase-2 variants; other
// variants ignore it)
void launch_contract(int nl, int ctc, bool miss, EpiMode mode, const ContractArgs& a, int n_col_groups,
                     int n_splits, hipStream_t st, int cpb = 1);
// workgroups of one contraction variant resident on the whole device at once
int contract_slots(int nl, int ctc, bool miss, EpiMode mode);
void launch_clear_tail(uint8_t* snp, int64_t m, int64_t stride, int64_t n, hipStream_t st);
// Chunk-tiled copy of a packed matrix of `rows` rows (a multiple of 32) and `stride`
// bytes per row (a multiple of 32): dst[r / 32][c][r % 32][32 B] = src[r][32 c .. 32 c + 32),
// so the 32 rows of a tile read by one contraction step for one 128-value chunk are
// 1 KiB of consecutive bytes.
void launch_tile_packed(const uint8_t* src, int64_t stride, int64_t rows, uint8_t* dst, hipStream_t st);
void launch_select_rows(const uint8_t* src, int64_t src_stride, int64_t m, const int64_t* rows, int64_t n_sel,
                        uint8_t* dst, int64_t dst_stride, hipStream_t st);
void launch_snp_stats(const uint8_t* snp, int64_t stride, int64_t m, int64_t n, int32_t* nobs, int64_t* sum,
                      int64_t* sumsq, hipStream_t st);
void launch_std_params(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                       int ddof, int completed, double* mean, double* inv_sd, hipStream_t st);
void launch_hwe_impute(uint8_t* snp, int64_t stride, int64_t m, int64_t n, const int32_t* nobs,
                       const int64_t* sum, uint64_t impute_seed, int64_t snp_offset, const int64_t* rows,
                       hipStream_t st);
void launch_transpose(const uint8_t* snp, int64_t snp_stride, int64_t m_pad, uint8_t* smp, int64_t smp_stride,
                      int64_t n_pad, hipStream_t st);
void launch_miss_count(const uint8_t* smp, int64_t stride, int64_t n, int32_t* counts, hipStream_t st);
// ptr[i] = sum_{t < i} counts[t] for i <= n (one workgroup)
void launch_scan_counts(const int32_t* counts, int64_t n, int64_t* ptr, hipStream_t st);
// Row-major limbs of the phase-2 operand for gathers by sample: xrow[((s * nch + h) * nl + l) * 64 + cc] =
// digit l of q = round(x * 2^shift) of column c = 64 h + cc of sample s (0 past the real columns), the
// same digits launch_quant_p2 swizzles for the contraction; rows s = 0..n, row n all zeros
void launch_xrow_limbs(const double* x, int64_t ldx, int64_t n, int n_bins, int v_count, int v_pad,
                       const int32_t* shift, int nl, int nch, int8_t* xrow, hipStream_t st);
// rptr[j * (parts + 1) + p] = first entry of SNP j's missing list (ptr/idx, sorted
// by sample) at a sample >= p * range, p = 0..parts
void launch_range_ptrs(const int64_t* ptr, const int32_t* idx, int64_t m, int parts, int64_t range, int64_t* rptr,
                       hipStream_t st);
// out[(j - r0) * ld_ms + c] += sum over the missing calls (n, j) of q[n][c], exact (int64 atomic
// adds of the per-sample-range sums; out zeroed by the caller), for SNPs j in [r0, r1), from the
// range pointers of launch_range_ptrs and the limb rows of launch_xrow_limbs
void launch_miss_sums(const int64_t* rptr, int parts, const int32_t* idx, int64_t r0, int64_t r1, const int8_t* xrow,
                      int64_t n, int nl, int nch, int64_t* out, int64_t ld_ms, hipStream_t st);
// out[k * ldo + i] += sum_{j in [s, e)} xt_ij^2 w[(j - s) + k ldw] (k < n_bins), xt the standardized
// working genotype minus C (R x) when p > 0 (rx[(j - s) ldrx + a] = (R x_j)_a, p <= 64); part holds
// 8 n ceil((e - s) / sq_rowsums_chunk(n, e - s)) doubles
int64_t sq_rowsums_chunk(int64_t n, int64_t L);
void launch_sq_rowsums(const uint8_t* smp, int64_t stride, int64_t n, int64_t s, int64_t e, const double* mean,
                       const double* inv_sd, const double* w, int64_t ldw, int n_bins, const double* C, int64_t ldc,
                       const double* rx, int64_t ldrx, int p, double* part, double* out, int64_t ldo, hipStream_t st);

void launch_miss_fill(const uint8_t* smp, int64_t stride, int64_t n, const int64_t* ptr, int32_t* idx,
                      hipStream_t st);
void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st);
void launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t st);
// sub > 0: probe columns in tiles of sub, each tile the reference's stream for V-tile v_start + t*sub
// v_start_dev (may be null): read the first probe id on the device instead of v_start
void launch_mt_rademacher(uint64_t root, const int64_t* blk_starts, int n_blocks, int64_t blk_end_all,
                          int v_start, int v_count, int8_t* z, int64_t ldz, hipStream_t st, int sub = 0,
                          const int32_t* v_start_dev = nullptr);
void launch_add_i32(int32_t* v, int32_t inc, hipStream_t st);
void launch_coef_p1_colmax(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                           int64_t k0, int64_t k1, int n_bins, int v_count, int v_pad, uint64_t* colmax_bits,
                           hipStream_t st);
void launch_coef_p1_quant(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                          int64_t k0, int64_t k1, int n_bins, int v_count, int v_pad, const int32_t* shift,
                          int nl, int ctc, int64_t coef_ctiles, int64_t kc0, int8_t* coef,
                          const double* mean, double* cst, double* cst_part, hipStream_t st);
// SNP slices of the Cst reduction (cst_part holds n_cols * kCstParts doubles)
constexpr int kCstParts = 128;
void launch_p1_finalize(const int64_t* y, int64_t ldy, int64_t n, int n_bins, int v_count, int v_pad,
                        const double* col_scale, const double* cst, const int64_t* miss_ptr,
                        const int32_t* miss_idx, int64_t k0, int64_t k1, const double* w, int64_t ldw,
                        const int8_t* z_i8, const double* z_f64, int64_t ldz, const int32_t* shift,
                        const double* mean, double* xz, int64_t ldxz, const uint8_t* col_init, hipStream_t st);
void launch_p1_miss_rec(const double* w, int64_t ldw, const int8_t* z, int64_t ldz, int64_t k0, int64_t k1,
                        int n_bins, int v_count, int v_pad, const int32_t* shift, const double* col_scale,
                        const double* mean, void* rec, hipStream_t st);
// bytes of launch_p1_miss_rec records for n_snps SNPs
int64_t p1_miss_rec_bytes(int64_t n_snps, int n_bins, int v_pad);
void launch_p1_finalize_z8(const int64_t* y, int64_t ldy, int64_t n, int n_bins, int v_count, int v_pad,
                           const double* col_scale, const double* cst, const int64_t* miss_ptr,
                           const int32_t* miss_idx, int64_t k0, int64_t k1, const void* rec, double* xz,
                           int64_t ldxz, const uint8_t* col_mode, int64_t range_bytes, hipStream_t st);
void launch_colmax_f64(const double* x, int64_t ldx, int64_t n, int n_cols, int v_count, int v_pad,
                       uint64_t* colmax_bits, hipStream_t st);
void launch_quant_p2(const double* x, int64_t ldx, int64_t n, int n_bins, int v_count, int v_pad,
                     const int32_t* shift, int nl, int ctc, int64_t coef_ctiles, int8_t* coef,
                     int64_t* col_sum, hipStream_t st);
void launch_z8_check(const int8_t* z, int64_t ldz, int64_t rows, int cols, unsigned int* bad, hipStream_t st);
void launch_col_sum_override(const double* sum, int n_bins, int v_count, int v_pad, const int32_t* shift,
                             double* col_scale, int64_t* col_sum, hipStream_t st);
void launch_shift_from_colmax(const uint64_t* colmax_bits, int n, int precision_bits, int32_t* shift,
                              double* col_scale, hipStream_t st);
void launch_p2_finalize(const double* s2, const double* s4, int64_t ld_s, int64_t r0, int64_t r1, int n_bins,
                        int tiles_per_bin, int v_count, const double* inv_left, double inv_denom2,
                        double extreme_warn, int drop_extreme, double* meansq, int64_t ld_meansq,
                        double* mc_agg, double* mc_per_snp, unsigned long long* guard_counts,
                        hipStream_t st);
void launch_coef_i8(const int8_t* z, int64_t ldz, int64_t k0, int64_t k1, int n_cols, int64_t coef_ctiles,
                    int8_t* coef, hipStream_t st);
void launch_quant_i32(const int32_t* u, int64_t ldu, int64_t n, int n_cols, int nl, int ctc,
                      int64_t coef_ctiles, int8_t* coef, hipStream_t st);
void launch_synth_markov(uint8_t* snp, int64_t stride, int64_t m, int64_t n, uint64_t seed, double maf_lo,
                         double maf_hi, double miss_rate, int block_len, double copy_prob, int iid,
                         hipStream_t st);
void launch_gram_band(const GramArgs& a, int64_t n_pairs, bool miss, hipStream_t st);
// G[k * p + q] = sum_i C[k ldc + i] C[q ldc + i] and csum[k] = sum_i C[k ldc + i], i < n
void launch_cov_gram(const double* C, int64_t ldc, int64_t n, int p, double* G, double* csum, hipStream_t st);
// Per-SNP terms of the projected, restandardized windowed panel for SNPs [s, e)
// (rows relative to s): u = C^T x_j, v = R x_j (row-major (L, p)), ||x_j||^2 from the
// integer moments and inv_sd; A_j = [G v_j - u_j, -v_j], B_j = [v_j, u_j], ybar_j =
// -(C^T 1) . v_j / n, t_j = 1 / sd of the projected column (1 when it has no variance)
void launch_win_cov_prep(const double* u, const double* v, int p, const double* G, const double* csum,
                         const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* inv_sd,
                         int completed, int64_t s, int64_t e, int64_t n, double* A, double* B, double* ybar,
                         double* tsc, hipStream_t st);
void launch_philox_rademacher(const uint64_t* keys, const int64_t* starts, const int64_t* lens, int n_blocks,
                              int v_count, int64_t max_len, int64_t row0, int8_t* z, int64_t ldz, hipStream_t st);
void launch_weighted_sumsq(const uint8_t* snp, int64_t stride, int64_t r0, int64_t r1, int64_t n, const double* mean,
                           const double* inv_sd, const double* W_sample_major, int ec, int n_w, double* out,
                           int64_t ldo, hipStream_t st);
// out[e * ldo + (j - r0)] = inv_sd_j^2 scale_e (mu_j^2 (wsum_e - me) - 2 mu_j ge + g2) for the
// three int64 plane sums (row-major [j - r0][e], leading dimension ld)
// xz[(bin*v_count + v) * ldxz + n] (=|+=) y[c][n] cs[c] - cst[c] + y2[c][n] cs2[c] (c = bin*v_pad + v)
void launch_p1_combine2(const int64_t* y, const int64_t* y2, int64_t ldy, int64_t n, int n_bins, int v_count,
                        int v_pad, const double* cs, const double* cst, const double* cs2, double* xz, int64_t ldxz,
                        const uint8_t* col_mode, hipStream_t st);
// out[i * n_bins + k] = mean[i] * w[i * n_bins + k] for i in [s, e)
void launch_scale_rows(const double* w, const double* mean, int64_t s, int64_t e, int n_bins, double* out,
                       hipStream_t st);
void launch_wss_finalize(const int64_t* ge, const int64_t* me, const int64_t* g2, int64_t ld, int64_t r0, int64_t r1,
                         int n_w, const double* mean, const double* inv_sd, const double* col_scale,
                         const int64_t* col_sum, double* out, int64_t ldo, hipStream_t st);
void launch_window_scale(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                         int completed, double* scale, hipStream_t st);
void launch_resvar_inv(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* mean,
                       const double* inv_sd, int64_t r0, int64_t r1, int64_t n, int completed, double eps,
                       double* inv_out, hipStream_t st);

// x_q <- x_q - A (B x_q) [- mean] for q_count FP64 columns; scratch T (p x q_count),
// S, mean (q_count) and a (p)
// S[q] = sum_{i < n} x[q ldx + i] (fixed-order reduction, one workgroup per column)
void launch_col_sums(const double* x, int64_t ldx, int64_t n, int q_count, double* S, hipStream_t st);
void launch_basis_coef(const double* x, int64_t ldx, int64_t n, int q_count, const double* B, int64_t ldb, int p,
                       double* T, int64_t ldt, double* s2, hipStream_t st);
void launch_project(double* x, int64_t ldx, int64_t n, int q_count, const double* A, int64_t lda, const double* B,
                    int64_t ldb, int p, bool center, double* T, double* S, double* a, double* mean, hipStream_t st);

// Records the calling thread's failure message (returned by gpe_last_error);
// entry points of other translation units report through it.
void set_last_error(const char* msg);

}  // namespace gpe
