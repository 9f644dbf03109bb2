/*
 * ORACLE — test infrastructure only.  Nothing in the product links this file;
 * only tests/ and bench.py's cpu_baseline leg load it (through
 * oracle/oracle.py), and only as the checker or as the timed CPU baseline.
 *
 * Plain-C + OpenMP restatement of the reference's Mailman products, the CPU
 * path the reference takes for at most 10 probes:
 *
 *   - segment size: floor(log3 N) - 1 clamped to [1, 8], 1 for N <= 27
 *     (src/native/common/genotype.cpp:186-200; the environment override is
 *     left out);
 *   - segment codes: SNPs j0..j0+s-1 of a sample form the base-3 number
 *     sum_t g_t 3^(s-1-t), SNP j0 the most significant digit
 *     (src/native/common/genotype.cpp:1020-1158 packs the same order that
 *     mailman.hpp decodes);
 *   - X.Z (post-multiply): a table of the 3^s code values built digit by
 *     digit from the last SNP, then one table row added per sample
 *     (src/native/common/mailman.hpp:213-270);
 *   - X^T U (pre-multiply): every sample's row added into its code's table
 *     row, then the table folded SNP by SNP into the per-SNP sums
 *     (src/native/common/mailman.hpp:17-78).
 *
 * Genotypes are the BED codes 00 -> 0, 10 -> 1, 11 -> 2 with missing calls
 * (01) as 0, and the products are exact int64: the integer products G.Z and
 * G^T U of the exact path, so the restatement is pinned by the dense integer
 * products (tests/test_oracle_mailman.py).
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int or_mailman_segment_size(int64_t n) {
    if (n <= 27) return 1;
    int seg = (int)floor(log((double)n) / log(3.0)) - 1;
    if (seg < 1) seg = 1;
    if (seg > 8) seg = 8;
    return seg;
}

static inline int bed_value(const uint8_t* row, int smp) {
    const int code = (row[smp >> 2] >> (2 * (smp & 3))) & 3;
    return code == 2 ? 1 : (code == 3 ? 2 : 0); /* 00 -> 0, 01 (missing) -> 0, 10 -> 1, 11 -> 2 */
}

static inline int segment_code(const uint8_t* bed, int64_t stride, int j0, int s, int smp) {
    int code = 0;
    for (int t = 0; t < s; ++t) code = code * 3 + bed_value(bed + (int64_t)(j0 + t) * stride, smp);
    return code;
}

/* out[smp * b + c] = sum_j g(j, smp) z[j * b + c], (n, b) row-major int64. */
void or_mailman_gz(const uint8_t* bed, int64_t stride, int n, int m, const int8_t* z, int b, int64_t* out,
                   int threads) {
    const int s = or_mailman_segment_size(n);
    const int n_seg = (m + s - 1) / s;
    memset(out, 0, (size_t)n * (size_t)b * sizeof(int64_t));
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
        int64_t tsize = 1;
        for (int i = 0; i < s; ++i) tsize *= 3;
        int64_t* table = (int64_t*)calloc((size_t)tsize * (size_t)b, sizeof(int64_t));
        int64_t* part = (int64_t*)calloc((size_t)n * (size_t)b, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int seg = 0; seg < n_seg; ++seg) {
            const int j0 = seg * s;
            const int sa = (m - j0) < s ? (m - j0) : s;
            /* table[code] = sum_t digit_t(code) z_(j0+t), built from the last SNP */
            for (int c = 0; c < b; ++c) table[c] = 0;
            int64_t prefix = 1;
            for (int t = sa - 1; t >= 0; --t) {
                const int8_t* zr = z + (int64_t)(j0 + t) * b;
                for (int64_t idx = 0; idx < prefix; ++idx) {
                    const int64_t* base = table + idx * b;
                    int64_t* one = table + (prefix + idx) * b;
                    int64_t* two = table + (2 * prefix + idx) * b;
                    for (int c = 0; c < b; ++c) {
                        one[c] = base[c] + zr[c];
                        two[c] = base[c] + 2 * (int64_t)zr[c];
                    }
                }
                prefix *= 3;
            }
            for (int smp = 0; smp < n; ++smp) {
                const int64_t* src = table + (int64_t)segment_code(bed, stride, j0, sa, smp) * b;
                int64_t* dst = part + (int64_t)smp * b;
                for (int c = 0; c < b; ++c) dst[c] += src[c];
            }
        }
#ifdef _OPENMP
#pragma omp critical
#endif
        {
            for (int64_t i = 0; i < (int64_t)n * b; ++i) out[i] += part[i];
        }
        free(part);
        free(table);
    }
}

/* out[j * b + c] = sum_smp g(j, smp) u[smp * b + c], (m, b) row-major int64. */
void or_mailman_gtu(const uint8_t* bed, int64_t stride, int n, int m, const int64_t* u, int b, int64_t* out,
                    int threads) {
    const int s = or_mailman_segment_size(n);
    const int n_seg = (m + s - 1) / s;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
#endif
    {
        int64_t tsize = 1;
        for (int i = 0; i < s; ++i) tsize *= 3;
        int64_t* table = (int64_t*)calloc((size_t)tsize * (size_t)b, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
        for (int seg = 0; seg < n_seg; ++seg) {
            const int j0 = seg * s;
            const int sa = (m - j0) < s ? (m - j0) : s;
            int64_t size = 1;
            for (int i = 0; i < sa; ++i) size *= 3;
            for (int smp = 0; smp < n; ++smp) {
                int64_t* dst = table + (int64_t)segment_code(bed, stride, j0, sa, smp) * b;
                const int64_t* src = u + (int64_t)smp * b;
                for (int c = 0; c < b; ++c) dst[c] += src[c];
            }
            /* fold: the leading digit of index + d (digit 1) or + 2d (digit 2) belongs to SNP j0+t */
            int64_t d = size;
            for (int t = 0; t < sa; ++t) {
                d /= 3;
                int64_t* o = out + (int64_t)(j0 + t) * b;
                for (int c = 0; c < b; ++c) o[c] = 0;
                for (int64_t idx = 0; idx < d; ++idx) {
                    int64_t* r0 = table + idx * b;
                    int64_t* r1 = table + (idx + d) * b;
                    int64_t* r2 = table + (idx + 2 * d) * b;
                    for (int c = 0; c < b; ++c) {
                        const int64_t one = r1[c], two = r2[c];
                        r1[c] = 0;
                        r2[c] = 0;
                        r0[c] += one + two;
                        o[c] += one + 2 * two;
                    }
                }
            }
            for (int c = 0; c < b; ++c) table[c] = 0;
        }
        free(table);
    }
}
