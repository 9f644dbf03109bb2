// HIP kernels of the MI355X genotype-probe engine (gfx950 / CDNA4).
//
// The engine computes the two genotype traversals of SUMMIT's stochastic
// LD-score estimator (src/native/gwldcore.cpp:1582-1861 phase 1 X*Z,
// :2329-2600 phase 2 X^T(X*Z)) on a packed 2-bit genotype matrix resident in
// HBM.  Every product is one kernel, k_contract: it streams packed rows with
// 16-byte loads, expands 2-bit codes to int8 in registers (genotype count and
// missing-call planes), and contracts them against int8 coefficient limbs on
// v_mfma_i32_32x32x32_i8 with exact int32 accumulation.  Real-valued
// coefficients (weighted probes, the FP64 X*Z fed to phase 2) are split into
// balanced base-256 int8 limbs with a per-column power-of-two scale, so each
// limb product is exact and the recombined int64 is the exact product of the
// quantized operand; centering, scaling and missing-call corrections are
// applied in FP64 in the epilogues.
#include "gpe_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <map>
#include <mutex>
#include <tuple>

namespace gpe {

typedef int32_t v4i __attribute__((ext_vector_type(4)));
typedef int32_t v16i __attribute__((ext_vector_type(16)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

// Unpacking 16 codes (one 32-bit word, code t in bits 2t..2t+1) into MFMA
// operand bytes.  Byte j of s[d] = (w >> 2d) & 0x03030303 holds code 4j + d,
// so the operand's byte position 4d + j carries code 4j + d: the contraction
// index is permuted (a 4x4 transpose) inside every group of 16.  The
// coefficient layout (swz_offset) applies the same permutation, and products
// of two unpacked operands see it on both sides, so every contraction sums the
// same pairs.  The code -> value maps are byte-table lookups with v_perm_b32
// (selector bytes 0..3 pick bytes of the table word):
//   genotype count 00 -> 0, 01 (missing) -> 0, 10 -> 1, 11 -> 2;
//   missing flag   01 -> 1, otherwise 0.
constexpr uint32_t kGenoTable = 0x02010000u;
constexpr uint32_t kMissTable = 0x00000100u;
// squared genotype count 00 -> 0, 01 (missing) -> 0, 10 -> 1, 11 -> 4
constexpr uint32_t kSqTable = 0x04010000u;

__device__ __forceinline__ void code_bytes(uint32_t w, uint32_t s[4]) {
    s[0] = w & 0x03030303u;
    s[1] = (w >> 2) & 0x03030303u;
    s[2] = (w >> 4) & 0x03030303u;
    s[3] = (w >> 6) & 0x03030303u;
}

// 16 codes -> 16 int8 genotype counts (missing -> 0) and 16 int8 missing flags.
__device__ __forceinline__ void unpack16(uint32_t w, v4i& g, v4i& m) {
    uint32_t s[4];
    code_bytes(w, s);
    g = v4i{(int32_t)__builtin_amdgcn_perm(0u, kGenoTable, s[0]), (int32_t)__builtin_amdgcn_perm(0u, kGenoTable, s[1]),
            (int32_t)__builtin_amdgcn_perm(0u, kGenoTable, s[2]), (int32_t)__builtin_amdgcn_perm(0u, kGenoTable, s[3])};
    m = v4i{(int32_t)__builtin_amdgcn_perm(0u, kMissTable, s[0]), (int32_t)__builtin_amdgcn_perm(0u, kMissTable, s[1]),
            (int32_t)__builtin_amdgcn_perm(0u, kMissTable, s[2]), (int32_t)__builtin_amdgcn_perm(0u, kMissTable, s[3])};
}

template <uint32_t TABLE = kGenoTable>
__device__ __forceinline__ void unpack16_g(uint32_t w, v4i& g) {
    uint32_t s[4];
    code_bytes(w, s);
    g = v4i{(int32_t)__builtin_amdgcn_perm(0u, TABLE, s[0]), (int32_t)__builtin_amdgcn_perm(0u, TABLE, s[1]),
            (int32_t)__builtin_amdgcn_perm(0u, TABLE, s[2]), (int32_t)__builtin_amdgcn_perm(0u, TABLE, s[3])};
}

// Offset of coefficient byte (c-tile tg, value index kk in [0,128), lane row lr in [0,32))
// inside one K chunk of the swizzled layout: [tile][step][lane][16 bytes].  Value kk
// sits in lane half kk/64, step (kk%64)/16 and, within its group of 16, at the
// transposed byte position the unpacked genotype operand uses.
__device__ __forceinline__ int64_t swz_offset(int64_t kl, int64_t coef_ctiles, int64_t tg, int kk, int lr) {
    const int h = kk >> 6;
    const int s = (kk & 63) >> 4;
    const int b = ((kk & 3) << 2) | ((kk >> 2) & 3);
    const int lane = h * 32 + lr;
    return (((kl * coef_ctiles + tg) * 4 + s) * 64 + lane) * 16 + b;
}

// Balanced base-256 digits of q, most significant first, into d[0..NL).
template <int NL>
__device__ __forceinline__ void balanced_digits(int64_t q, int8_t* d) {
#pragma unroll
    for (int l = NL - 1; l >= 0; --l) {
        const int64_t r = ((q + 128) & 255) - 128;
        d[l] = (int8_t)r;
        q = (q - r) / 256;
    }
}

// round(v 2^sh) as an integer; 0 for a non-finite or out-of-range value (the
// column's scale is NaN then, so the value never reaches an output).
__device__ __forceinline__ int64_t quant_rn(double v, int sh) {
    const double t = ldexp(v, sh);
    return fabs(t) < 4.0e18 ? __double2ll_rn(t) : 0;
}

__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ---------------------------------------------------------------------------
// Packed contraction on int8 MFMA.
//
//   out[c][r] = sum_k A[c][k] * plane(P[r][k])
//
// Workgroup: 4 waves.  Each wave owns 32 packed rows (MISS: genotype and
// missing planes of the same rows) or 64 rows (two genotype tiles), and all
// CT = NL*CTC coefficient tiles of one column group.  Per K chunk of 128
// values the workgroup's CT*4 KiB of swizzled coefficients arrive in LDS by
// direct global->LDS loads (three buffers, issued two chunks ahead, no
// registers), and each lane loads 16 packed bytes (64 codes) of its row into
// registers, also two chunks ahead.
// MFMA lane layout used here: operand lane l holds row (l & 31) and 16
// consecutive K values of half (l >> 5); result register e of lane l holds
// (row 8*(e>>2) + 4*(l>>5) + (e&3), column l & 31).
// ---------------------------------------------------------------------------
// column groups per band of the contraction's workgroup order (GxE at N=100,000,
// M=500,000, 2,048 phase-2 columns: 0.83 s with 4, 0.82 s with 2, 0.88 s with 8,
// 0.91 s unbanded; partitioned, 800 columns: 0.117-0.119 s with 4, 0.119-0.123 s with 2)
constexpr int kGroupBand = 4;

constexpr uint32_t waitcnt_vm(uint32_t n) {  // s_waitcnt vmcnt(n), other counters unconstrained
    return (n & 15u) | ((n >> 4) << 14) | (7u << 4) | (15u << 8);
}

// One K chunk of a column group's coefficients (CT x 4 KiB, contiguous at src_chunk)
// into LDS at dst by direct buffer loads (MUBUF with the LDS bit; the 256 threads take
// 16 bytes each per 4 KiB).  The buffer form, not global_load_lds: the compiler counts
// a global_load_lds as a FLAT access that may touch LDS, and while one is in flight it
// turns every wait for an LDS operand read into lgkmcnt(0) -- every read issued since
// must land -- which serializes the operand reads of the MFMA steps.  Buffer loads
// are counted by vmcnt only, so those waits stay exact.
template <int CT>
__device__ __forceinline__ void coef_chunk_to_lds(const int8_t* src_chunk, int8_t* dst, int tid, int wave) {
    const __amdgpu_buffer_rsrc_t rsrc =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<int8_t*>(src_chunk), (short)0, CT * 4096, 0x00020000);
#pragma unroll
    for (int t = 0; t < CT; ++t)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(rsrc, (__attribute__((address_space(3))) void*)(dst + (t * 256 + wave * 64) * 16),
                                                 16, (t * 256 + tid) * 16, 0, 0, 0);
}

// The four MFMA steps of one 128-value chunk whose coefficients sit at L:
// operands op0/op1 of step 0 are already unpacked; each step unpacks the next
// step's operands first (from ga/gb, after step 3 from the words na/nb of the
// following chunk) so the unpack interleaves with this step's MFMAs.
template <int CT, bool MISS, uint32_t TBL>
__device__ __forceinline__ void chunk_steps(const int8_t* L, int lane, const v4i& ga, const v4i& gb, uint32_t na,
                                            uint32_t nb, v4i& op0, v4i& op1, v16i (&acc)[CT][2]) {
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const uint32_t wn0 = s < 3 ? (uint32_t)ga[s + 1] : na;
        v4i nx0, nx1;
        if (MISS) {
            unpack16(wn0, nx0, nx1);
        } else {
            const uint32_t wn1 = s < 3 ? (uint32_t)gb[s + 1] : nb;
            unpack16_g<TBL>(wn0, nx0);
            unpack16_g<TBL>(wn1, nx1);
        }
#pragma unroll
        for (int t = 0; t < CT; ++t) {
            const v4i av = *reinterpret_cast<const v4i*>(L + ((t * 4 + s) * 64 + lane) * 16);
            acc[t][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, op0, acc[t][0], 0, 0, 0);
            acc[t][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, op1, acc[t][1], 0, 0, 0);
        }
        op0 = nx0;
        op1 = nx1;
    }
}

// Epilogues of the packed contraction, per wave: acc[l*CTC + ctc][bt] holds limb l of
// c-tile cg*CTC + ctc for row tile bt (rows rbase + 32*bt + lr; MISS: tile 1 is the
// missing plane of the same rows).
template <int NL, int CTC, bool MISS, int EPI>
__device__ __forceinline__ void contract_epilogue(const ContractArgs& a, v16i (&acc)[NL * CTC][2], int cg,
                                                  int64_t rbase, int lr, int half) {
    if constexpr (EPI == EPI_I32_ATOMIC) {
        static_assert(NL == 1, "int32 output needs one limb");
#pragma unroll
        for (int ctc = 0; ctc < CTC; ++ctc) {
            const int ct = cg * CTC + ctc;
            if (ct >= a.ctiles_used) continue;
#pragma unroll
            for (int bt = 0; bt < 2; ++bt) {
                const int64_t r = rbase + (MISS ? 0 : bt * 32) + lr;
                if (MISS && bt == 1) continue;
                if (r >= a.r1) continue;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const int c = ct * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                    if (c < a.cols_out) atomicAdd(&a.out_i32[(int64_t)c * a.ldo + (r - a.out_row0)], acc[ctc][bt][e]);
                }
            }
        }
    } else if constexpr (EPI == EPI_I64_2P) {
        static_assert(MISS, "two-plane output needs the missing plane");
#pragma unroll
        for (int ctc = 0; ctc < CTC; ++ctc) {
            const int ct = cg * CTC + ctc;
            if (ct >= a.ctiles_used) continue;
            const int64_t r = rbase + lr;
            if (r >= a.r1) continue;
#pragma unroll
            for (int e = 0; e < 16; ++e) {
                int64_t ge = 0, me = 0;
#pragma unroll
                for (int l = 0; l < NL; ++l) {
                    ge = ge * 256 + (int64_t)acc[l * CTC + ctc][0][e];
                    me = me * 256 + (int64_t)acc[l * CTC + ctc][1][e];
                }
                const int c = ct * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                if (c >= a.cols_out) continue;
                a.out_i64[(r - a.out_row0) * a.ldo + c] = ge;
                a.out2_i64[(r - a.out_row0) * a.ldo + c] = me;
            }
        }
    } else if constexpr (EPI == EPI_I64_ATOMIC || EPI == EPI_I64_ROWMAJOR || EPI == EPI_I64_SQ ||
                         EPI == EPI_I64_ATOMIC_Q) {
#pragma unroll
        for (int ctc = 0; ctc < CTC; ++ctc) {
            const int ct = cg * CTC + ctc;
            if (ct >= a.ctiles_used) continue;
#pragma unroll
            for (int bt = 0; bt < 2; ++bt) {
                if (MISS && bt == 1) continue;
                const int64_t r = rbase + (MISS ? 0 : bt * 32) + lr;
                if (r >= a.r1) continue;
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    int64_t v = 0;
#pragma unroll
                    for (int l = 0; l < NL; ++l) v = v * 256 + (int64_t)acc[l * CTC + ctc][bt][e];
                    const int c = ct * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                    if (c >= a.cols_out) continue;
                    if (EPI == EPI_I64_ATOMIC || EPI == EPI_I64_ATOMIC_Q)
                        atomicAdd(reinterpret_cast<unsigned long long*>(
                                      &a.out_i64[(int64_t)c * a.ldo + (r - a.out_row0)]),
                                  (unsigned long long)v);
                    else
                        a.out_i64[(r - a.out_row0) * a.ldo + c] = v;
                }
            }
        }
    } else if constexpr (EPI == EPI_WORK) {
#pragma unroll
        for (int ctc = 0; ctc < CTC; ++ctc) {
            const int ct = cg * CTC + ctc;
#pragma unroll
            for (int bt = 0; bt < (MISS ? 1 : 2); ++bt) {
                const int64_t r = rbase + (MISS ? 0 : bt * 32) + lr;
                if (r >= a.r1 || ct >= a.ctiles_used) continue;
                const double mu = a.mean[r];
                const double isd = a.inv_sd[r];
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    int64_t ge = 0, me = 0;
#pragma unroll
                    for (int l = 0; l < NL; ++l) {
                        ge = ge * 256 + (int64_t)acc[l * CTC + ctc][bt][e];
                        if (MISS) me = me * 256 + (int64_t)acc[l * CTC + ctc][1][e];
                    }
                    const int c = ct * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                    if (c >= a.cols_out) continue;
                    const double obs_sum = (double)(a.col_sum[c] - me);
                    a.out_f64[(r - a.out_row0) * a.ldo + c] = isd * a.col_scale[c] * ((double)ge - mu * obs_sum);
                }
            }
        }
    } else {  // EPI_P2, EPI_P2X
#pragma unroll
        for (int ctc = 0; ctc < CTC; ++ctc) {
            const int ct = cg * CTC + ctc;
#pragma unroll
            for (int bt = 0; bt < (MISS ? 1 : 2); ++bt) {
                const int64_t r = rbase + (MISS ? 0 : bt * 32) + lr;
                const bool live = r < a.r1 && ct < a.ctiles_used;
                double s2 = 0.0, s4 = 0.0;
                if (live) {
                    const double mu = a.mean[r];
                    const double isd = a.inv_sd[r];
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        int64_t ge = 0, me = 0;
#pragma unroll
                        for (int l = 0; l < NL; ++l) {
                            ge = ge * 256 + (int64_t)acc[l * CTC + ctc][bt][e];
                            if (MISS) me = me * 256 + (int64_t)acc[l * CTC + ctc][1][e];
                        }
                        const int c = ct * 32 + 8 * (e >> 2) + 4 * half + (e & 3);
                        if constexpr (EPI == EPI_P2X) me = a.msum[(r - a.r0) * a.ld_ms + c];
                        const double obs_sum = (double)(a.col_sum[c] - me);
                        const double w = isd * a.col_scale[c] * ((double)ge - mu * obs_sum);
                        const double w2 = w * w;
                        s2 += w2;
                        s4 += w2 * w2;
                    }
                }
                s2 += __shfl_xor(s2, 32, 64);
                s4 += __shfl_xor(s4, 32, 64);
                if (live && half == 0) {
                    a.s2[r * a.ld_s + ct] = s2;
                    if (a.s4) a.s4[r * a.ld_s + ct] = s4;
                }
            }
        }
    }
}

// CPB = chunks per barrier: 1 -> a three-slot LDS ring of one chunk per slot,
// loads issued two chunks ahead, one barrier per chunk; 2 -> a two-slot ring of
// two chunks per slot, loads issued one slot (two chunks) ahead, one barrier
// per two chunks (needs CT <= 5 to keep two workgroups per CU in 160 KiB of LDS).
template <int NL, int CTC, bool MISS, int EPI, int CPB = 1>
__global__ __launch_bounds__(256, (NL * CTC <= 2) ? 4 : 2) void k_contract(ContractArgs a) {
    constexpr int CT = NL * CTC;
    constexpr int RPW = MISS ? 32 : 64;
    constexpr int RPB = 4 * RPW;
    constexpr int G = MISS ? 1 : 2;  // packed-row loads per lane per chunk
    constexpr uint32_t TBL = EPI == EPI_I64_SQ ? kSqTable : (EPI == EPI_I64_ATOMIC_Q ? kMissTable : kGenoTable);
    static_assert(CPB == 1 || CPB == 2, "one or two chunks per barrier");
    constexpr int SLOTS = CPB == 1 ? 3 : 2;
    __shared__ __attribute__((aligned(16))) int8_t lds[SLOTS][CPB * CT * 4096];

    const int tid = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int lane = tid & 63;
    const int half = lane >> 5;
    const int lr = lane & 31;
    // Workgroups are dispatched in flattened (x, y, z) order and dealt round-robin
    // to the XCDs, each with its own L2.  The flattened index is remapped so that
    // every XCD runs a contiguous range of (split, row block, column group) tiles,
    // in bands of at most kGroupBand column groups: within a band the groups of a
    // row block run together (sharing its packed rows) and so do neighbouring row
    // blocks (sharing the band's coefficient chunks), so the ~64 workgroups an XCD
    // holds span about 64 / kGroupBand row blocks whatever the column count.
    const int n_groups = (int)gridDim.y;
    const int64_t gx = gridDim.x;
    const int64_t total = gx * n_groups * gridDim.z;
    const int64_t d = blockIdx.x + gx * (blockIdx.y + (int64_t)n_groups * blockIdx.z);
    const int64_t per_xcd = total / kXcds;
    const int64_t lin = d < per_xcd * kXcds ? (d % kXcds) * per_xcd + d / kXcds : d;
    const int64_t per_split = gx * n_groups;
    const int64_t split = lin / per_split;
    const int64_t t = lin - split * per_split;
    const int gband = n_groups < kGroupBand ? n_groups : kGroupBand;
    const int band = (int)(t / (gx * gband));
    const int gb_eff = n_groups - band * gband < gband ? n_groups - band * gband : gband;
    const int64_t rem = t - (int64_t)band * gx * gband;
    const int cg = band * gband + (int)(rem % gb_eff);
    const int64_t rblk = rem / gb_eff;
    const int64_t rbase = a.r0 + rblk * RPB + (int64_t)wave * RPW;
    const int64_t kc_begin = a.kc0 + split * a.kc_per_split;
    const int64_t kc_end = kc_begin + a.kc_per_split < a.kc1 ? kc_begin + a.kc_per_split : a.kc1;

    v16i acc[CT][2];
#pragma unroll
    for (int t = 0; t < CT; ++t) {
        acc[t][0] = v16i{};
        acc[t][1] = v16i{};
    }

    // row r of the packed operand: row-major (a.tiled == 0) at r * stride, chunk kc at
    // + 32 kc; chunk-tiled (a.tiled == 1, [r / 32][chunk][r % 32][32 B]) at
    // (r / 32) * 32 * stride + (r % 32) * 32, chunk kc at + 1024 kc.  Either way the
    // next 32-row tile is 32 * stride further.
    const int64_t row = rbase + lr;
    const int64_t cpitch = a.tiled ? 1024 : 32;
    const uint8_t* prow0 = a.packed + (row >> 5) * (32 * a.stride) + (row & 31) * (a.tiled ? 32 : a.stride) + half * 16;
    const uint8_t* prow1 = MISS ? prow0 : prow0 + 32 * a.stride;
    const int8_t* coef_grp = a.coef + (int64_t)cg * CT * 4096;
    const int64_t coef_chunk_bytes = a.coef_ctiles * 4096;

    if constexpr (CPB == 1) {
        // packed words of chunks kc (g0), kc+1 (g1); plain unrolled code, no lambdas,
        // so these stay in VGPRs.  Prefetches past the end re-read the last chunk
        // (in bounds, never used).
        v4i g0a = v4i{}, g0b = v4i{}, g1a = v4i{}, g1b = v4i{};
        if (kc_begin < kc_end) {
            const int64_t k1 = kc_begin + 1 < kc_end ? kc_begin + 1 : kc_end - 1;
            coef_chunk_to_lds<CT>(coef_grp + (kc_begin - a.kc0) * coef_chunk_bytes, lds[0], tid, wave);
            g0a = *reinterpret_cast<const v4i*>(prow0 + kc_begin * cpitch);
            if (!MISS) g0b = *reinterpret_cast<const v4i*>(prow1 + kc_begin * cpitch);
            coef_chunk_to_lds<CT>(coef_grp + (k1 - a.kc0) * coef_chunk_bytes, lds[1], tid, wave);
            g1a = *reinterpret_cast<const v4i*>(prow0 + k1 * cpitch);
            if (!MISS) g1b = *reinterpret_cast<const v4i*>(prow1 + k1 * cpitch);
        }
        // operands of step 0 of the first chunk; later ones are unpacked one step ahead
        v4i op0, op1;
        if (MISS) {
            unpack16((uint32_t)g0a[0], op0, op1);
        } else {
            unpack16_g<TBL>((uint32_t)g0a[0], op0);
            unpack16_g<TBL>((uint32_t)g0b[0], op1);
        }
        int buf = 0;  // LDS buffer of chunk kc
        for (int64_t kc = kc_begin; kc < kc_end; ++kc) {
            // chunk kc's coefficients (and packed words) have landed once only the
            // CT + G loads issued for chunk kc+1 may still be in flight; the barrier
            // then publishes every wave's part of the chunk
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(CT + G));
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            // chunk kc+2 into the buffer read two chunks ago (k2 passes through an empty
            // asm so these loads cannot be hoisted above the wait and barrier)
            int64_t k2 = kc + 2 < kc_end ? kc + 2 : kc_end - 1;
            asm volatile("" : "+s"(k2));
            const int buf2 = buf == 0 ? 2 : buf - 1;
            coef_chunk_to_lds<CT>(coef_grp + (k2 - a.kc0) * coef_chunk_bytes, lds[buf2], tid, wave);
            const v4i g2a = *reinterpret_cast<const v4i*>(prow0 + k2 * cpitch);
            const v4i g2b = MISS ? v4i{} : *reinterpret_cast<const v4i*>(prow1 + k2 * cpitch);
            const int8_t* L = lds[buf];
    #pragma unroll
            for (int s = 0; s < 4; ++s) {
                // unpack the operands of the next step (step 0 of the next chunk after
                // step 3) before this step's MFMAs, so the two are independent and interleave
                const uint32_t wn0 = s < 3 ? (uint32_t)g0a[s + 1] : (uint32_t)g1a[0];
                v4i nx0, nx1;
                if (MISS) {
                    unpack16(wn0, nx0, nx1);
                } else {
                    const uint32_t wn1 = s < 3 ? (uint32_t)g0b[s + 1] : (uint32_t)g1b[0];
                    unpack16_g<TBL>(wn0, nx0);
                    unpack16_g<TBL>(wn1, nx1);
                }
    #pragma unroll
                for (int t = 0; t < CT; ++t) {
                    const v4i av = *reinterpret_cast<const v4i*>(L + ((t * 4 + s) * 64 + lane) * 16);
                    acc[t][0] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, op0, acc[t][0], 0, 0, 0);
                    acc[t][1] = __builtin_amdgcn_mfma_i32_32x32x32_i8(av, op1, acc[t][1], 0, 0, 0);
                }
                op0 = nx0;
                op1 = nx1;
            }
            g0a = g1a;
            g0b = g1b;
            g1a = g2a;
            g1b = g2b;
            buf = buf == 2 ? 0 : buf + 1;
        }
    } else {
        // super chunk = chunks kc, kc+1 in one slot; words of the current pair in
        // g0/g1, of the next pair in g2/g3 (loaded one slot ahead)
        constexpr int64_t kSlot = (int64_t)CT * 4096;
        v4i g0a = v4i{}, g0b = v4i{}, g1a = v4i{}, g1b = v4i{};
        if (kc_begin < kc_end) {
            const int64_t c1 = kc_begin + 1 < kc_end ? kc_begin + 1 : kc_end - 1;
            coef_chunk_to_lds<CT>(coef_grp + (kc_begin - a.kc0) * coef_chunk_bytes, lds[0], tid, wave);
            coef_chunk_to_lds<CT>(coef_grp + (c1 - a.kc0) * coef_chunk_bytes, lds[0] + kSlot, tid, wave);
            g0a = *reinterpret_cast<const v4i*>(prow0 + kc_begin * cpitch);
            if (!MISS) g0b = *reinterpret_cast<const v4i*>(prow1 + kc_begin * cpitch);
            g1a = *reinterpret_cast<const v4i*>(prow0 + c1 * cpitch);
            if (!MISS) g1b = *reinterpret_cast<const v4i*>(prow1 + c1 * cpitch);
        }
        v4i op0, op1;
        if (MISS) {
            unpack16((uint32_t)g0a[0], op0, op1);
        } else {
            unpack16_g<TBL>((uint32_t)g0a[0], op0);
            unpack16_g<TBL>((uint32_t)g0b[0], op1);
        }
        int buf = 0;
        for (int64_t kc = kc_begin; kc < kc_end; kc += 2) {
            // everything issued for this pair (coefficients into LDS by every wave,
            // this lane's packed words) has landed; the barrier publishes the slot
            // and frees the other one, read during the previous pair
            __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
            asm volatile("" ::: "memory");
            __builtin_amdgcn_s_barrier();
            asm volatile("" ::: "memory");
            int64_t n0 = kc + 2 < kc_end ? kc + 2 : kc_end - 1;
            int64_t n1 = kc + 3 < kc_end ? kc + 3 : kc_end - 1;
            asm volatile("" : "+s"(n0));
            asm volatile("" : "+s"(n1));
            int8_t* nxt = lds[buf ^ 1];
            coef_chunk_to_lds<CT>(coef_grp + (n0 - a.kc0) * coef_chunk_bytes, nxt, tid, wave);
            coef_chunk_to_lds<CT>(coef_grp + (n1 - a.kc0) * coef_chunk_bytes, nxt + kSlot, tid, wave);
            const v4i g2a = *reinterpret_cast<const v4i*>(prow0 + n0 * cpitch);
            const v4i g2b = MISS ? v4i{} : *reinterpret_cast<const v4i*>(prow1 + n0 * cpitch);
            const v4i g3a = *reinterpret_cast<const v4i*>(prow0 + n1 * cpitch);
            const v4i g3b = MISS ? v4i{} : *reinterpret_cast<const v4i*>(prow1 + n1 * cpitch);
            const int8_t* L = lds[buf];
            chunk_steps<CT, MISS, TBL>(L, lane, g0a, g0b, (uint32_t)g1a[0], (uint32_t)g1b[0], op0, op1, acc);
            if (kc + 1 < kc_end)
                chunk_steps<CT, MISS, TBL>(L + kSlot, lane, g1a, g1b, (uint32_t)g2a[0], (uint32_t)g2b[0], op0, op1,
                                           acc);
            g0a = g2a;
            g0b = g2b;
            g1a = g3a;
            g1b = g3b;
            buf ^= 1;
        }
    }
    // no LDS-load may still be writing when the workgroup ends
    __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
    contract_epilogue<NL, CTC, MISS, EPI>(a, acc, cg, rbase, lr, half);
}



template <int NL, int CTC, bool MISS, int EPI, int CPB = 1>
static void launch_one(const ContractArgs& a, int n_col_groups, int n_splits, hipStream_t st) {
    constexpr int RPB = 4 * (MISS ? 32 : 64);
    const int64_t rows = a.r1 - a.r0;
    if (rows <= 0 || a.kc1 <= a.kc0) return;
    dim3 grid((unsigned)ceil_div(rows, RPB), (unsigned)n_col_groups, (unsigned)n_splits);
    hipLaunchKernelGGL((k_contract<NL, CTC, MISS, EPI, CPB>), grid, dim3(256), 0, st, a);
    GPE_HIP(hipGetLastError());
}

void launch_contract(int nl, int ctc, bool miss, EpiMode mode, const ContractArgs& a, int n_col_groups,
                     int n_splits, hipStream_t st, int cpb) {
    // two chunks per barrier: the phase-1 and phase-2 variants of the five-limb path
    if (cpb == 2) {
        if (nl == 5 && ctc == 1 && !miss && mode == EPI_I64_ATOMIC) {
            launch_one<5, 1, false, EPI_I64_ATOMIC, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 5 && ctc == 1 && miss && mode == EPI_P2) {
            launch_one<5, 1, true, EPI_P2, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 5 && ctc == 1 && !miss && mode == EPI_P2) {
            launch_one<5, 1, false, EPI_P2, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 5 && ctc == 1 && !miss && mode == EPI_P2X) {
            launch_one<5, 1, false, EPI_P2X, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 5 && ctc == 1 && !miss && mode == EPI_I64_ATOMIC_Q) {
            launch_one<5, 1, false, EPI_I64_ATOMIC_Q, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 1 && ctc == 2 && !miss && mode == EPI_I32_ATOMIC) {
            launch_one<1, 2, false, EPI_I32_ATOMIC, 2>(a, n_col_groups, n_splits, st);
            return;
        }
        if (nl == 3 && ctc == 1 && !miss && mode == EPI_I64_ROWMAJOR) {
            launch_one<3, 1, false, EPI_I64_ROWMAJOR, 2>(a, n_col_groups, n_splits, st);
            return;
        }
    }
#define GPE_CASE(NL_, CTC_, MISS_, EPI_)                                                          \
    if (nl == NL_ && ctc == CTC_ && miss == MISS_ && mode == EPI_) {                              \
        launch_one<NL_, CTC_, MISS_, EPI_>(a, n_col_groups, n_splits, st);                        \
        return;                                                                                   \
    }
    // exact integer products
    GPE_CASE(1, 2, false, EPI_I32_ATOMIC)
    GPE_CASE(3, 1, false, EPI_I64_ROWMAJOR)
    // weighted phase 1 (sample-major rows, int64 partial sums)
    GPE_CASE(5, 1, false, EPI_I64_ATOMIC)
    GPE_CASE(6, 1, false, EPI_I64_ATOMIC)
    // missing-flag plane of phase 1 with real-valued probes
    GPE_CASE(5, 1, false, EPI_I64_ATOMIC_Q)
    GPE_CASE(6, 1, false, EPI_I64_ATOMIC_Q)
    // plane sums of weighted sums of squares (genotype + missing, squared genotype)
    GPE_CASE(6, 1, true, EPI_I64_2P)
    GPE_CASE(6, 1, false, EPI_I64_ROWMAJOR)
    GPE_CASE(6, 1, false, EPI_I64_SQ)
    // weighted phase 2
    GPE_CASE(5, 1, false, EPI_P2)
    GPE_CASE(5, 1, true, EPI_P2)
    GPE_CASE(6, 1, false, EPI_P2)
    GPE_CASE(6, 1, true, EPI_P2)
    GPE_CASE(5, 1, false, EPI_P2X)
    GPE_CASE(6, 1, false, EPI_P2X)
    // dense standardized X^T V
    GPE_CASE(5, 1, false, EPI_WORK)
    GPE_CASE(5, 1, true, EPI_WORK)
    GPE_CASE(6, 1, false, EPI_WORK)
    GPE_CASE(6, 1, true, EPI_WORK)
#undef GPE_CASE
    throw Error("no packed-contraction variant for nl=" + std::to_string(nl) + " ctc=" + std::to_string(ctc) +
                " miss=" + std::to_string((int)miss) + " mode=" + std::to_string((int)mode));
}

template <int NL, int CTC, bool MISS, int EPI>
static int blocks_per_cu_one() {
    int nb = 0;
    GPE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void*)k_contract<NL, CTC, MISS, EPI>, 256, 0));
    return nb;
}

int contract_slots(int nl, int ctc, bool miss, EpiMode mode) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, bool, int>, int> cache;
    int dev = 0;
    GPE_HIP(hipGetDevice(&dev));
    const auto key = std::make_tuple(dev, nl, ctc, miss, (int)mode);
    std::lock_guard<std::mutex> lock(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
    int cus = 0;
    GPE_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    int nb = 0;
#define GPE_CASE(NL_, CTC_, MISS_, EPI_) \
    if (nl == NL_ && ctc == CTC_ && miss == MISS_ && mode == EPI_) nb = blocks_per_cu_one<NL_, CTC_, MISS_, EPI_>();
    GPE_CASE(1, 2, false, EPI_I32_ATOMIC)
    GPE_CASE(3, 1, false, EPI_I64_ROWMAJOR)
    GPE_CASE(5, 1, false, EPI_I64_ATOMIC)
    GPE_CASE(6, 1, false, EPI_I64_ATOMIC)
#undef GPE_CASE
    const int slots = cus * (nb > 0 ? nb : 1);
    cache[key] = slots;
    return slots;
}

// ---------------------------------------------------------------------------
// Layout / statistics kernels
// ---------------------------------------------------------------------------
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, int64_t n, uint64_t v) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = v;
}
void launch_fill_u8(uint8_t* p, int64_t n, uint8_t v, hipStream_t st) {
    if (n <= 0) return;
    const int64_t blocks = ceil_div(n, 256) < 65536 ? ceil_div(n, 256) : 65536;
    hipLaunchKernelGGL(k_fill_u8, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    GPE_HIP(hipGetLastError());
}
void launch_fill_u64(uint64_t* p, int64_t n, uint64_t v, hipStream_t st) {
    if (n <= 0) return;
    const int64_t blocks = ceil_div(n, 256) < 65536 ? ceil_div(n, 256) : 65536;
    hipLaunchKernelGGL(k_fill_u64, dim3((unsigned)blocks), dim3(256), 0, st, p, n, v);
    GPE_HIP(hipGetLastError());
}

// One thread per 16 bytes of the tiled copy: writes are consecutive, and the 64 lanes
// of a wave read the 32-byte chunk pieces of 32 rows.
__global__ void k_tile_packed(const uint8_t* __restrict__ src, int64_t stride, int64_t pieces,
                              uint8_t* __restrict__ dst) {
    const int64_t chunks = stride / 32;
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < pieces; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t off = t * 16;
        const int64_t tile = off / (chunks * 1024);
        const int64_t rem = off - tile * chunks * 1024;
        const int64_t c = rem >> 10;
        const int64_t lr = (rem & 1023) >> 5;
        const int64_t b = rem & 31;
        *reinterpret_cast<v4u*>(dst + off) =
            *reinterpret_cast<const v4u*>(src + (tile * 32 + lr) * stride + c * 32 + b);
    }
}
void launch_tile_packed(const uint8_t* src, int64_t stride, int64_t rows, uint8_t* dst, hipStream_t st) {
    if (rows <= 0) return;
    if (rows % 32 != 0 || stride % 32 != 0) throw Error("tiled copy needs rows and stride in multiples of 32");
    const int64_t pieces = rows * stride / 16;
    const int64_t blocks = std::min<int64_t>(ceil_div(pieces, 256), (int64_t)1 << 20);
    hipLaunchKernelGGL(k_tile_packed, dim3((unsigned)blocks), dim3(256), 0, st, src, stride, pieces, dst);
    GPE_HIP(hipGetLastError());
}

// Zero the codes of a SNP-major row past the last sample (padding must read as 00).
__global__ void k_clear_tail(uint8_t* snp, int64_t m, int64_t stride, int64_t n) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    uint8_t* row = snp + j * stride;
    const int64_t full = n / 4;
    const int rem = (int)(n % 4);
    int64_t b = full;
    if (rem) {
        row[b] = (uint8_t)(row[b] & ((1u << (2 * rem)) - 1u));
        ++b;
    }
    for (; b < stride; ++b) row[b] = 0;
}
void launch_clear_tail(uint8_t* snp, int64_t m, int64_t stride, int64_t n, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_clear_tail, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, snp, m, stride, n);
    GPE_HIP(hipGetLastError());
}

// Gather selected samples (sorted global row ids) into a compact SNP-major matrix.
__global__ void k_select_rows(const uint8_t* src, int64_t src_stride, int64_t m, const int64_t* rows,
                              int64_t n_sel, uint8_t* dst, int64_t dst_stride) {
    const int64_t nb = ceil_div(n_sel, 4);
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y;
    if (idx >= nb || j >= m) return;
    const uint8_t* s = src + j * src_stride;
    uint32_t out = 0;
    for (int t = 0; t < 4; ++t) {
        const int64_t n = idx * 4 + t;
        if (n >= n_sel) break;
        const int64_t g = rows[n];
        const uint32_t code = (s[g >> 2] >> (2 * (g & 3))) & 3u;
        out |= code << (2 * t);
    }
    dst[j * dst_stride + idx] = (uint8_t)out;
}
void launch_select_rows(const uint8_t* src, int64_t src_stride, int64_t m, const int64_t* rows, int64_t n_sel,
                        uint8_t* dst, int64_t dst_stride, hipStream_t st) {
    if (m <= 0 || n_sel <= 0) return;
    for (int64_t j0 = 0; j0 < m; j0 += 65535) {
        const int64_t mm = (m - j0) < 65535 ? (m - j0) : 65535;
        hipLaunchKernelGGL(k_select_rows, dim3((unsigned)ceil_div(ceil_div(n_sel, 4), 256), (unsigned)mm), dim3(256),
                           0, st, src + j0 * src_stride, src_stride, mm, rows, n_sel, dst + j0 * dst_stride,
                           dst_stride);
        GPE_HIP(hipGetLastError());
    }
}

// Per-SNP observed count, allele sum and sum of squares (genotype.cpp:474-574).
__global__ void k_snp_stats(const uint8_t* snp, int64_t stride, int64_t m, int64_t n, int32_t* nobs,
                            int64_t* sum, int64_t* sumsq) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= m) return;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(snp + j * stride);
    const int64_t words = stride / 4;
    int64_t nmiss = 0, c1 = 0, c2 = 0;
    for (int64_t w = lane; w < words; w += 64) {
        const uint32_t x = row[w];
        const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
        nmiss += __popc(lo & ~hi);
        c1 += __popc(hi & ~lo);
        c2 += __popc(hi & lo);
    }
    nmiss = wave_sum_i64(nmiss);
    c1 = wave_sum_i64(c1);
    c2 = wave_sum_i64(c2);
    if (lane == 0) {
        nobs[j] = (int32_t)(n - nmiss);
        sum[j] = c1 + 2 * c2;
        sumsq[j] = c1 + 4 * c2;
    }
}
void launch_snp_stats(const uint8_t* snp, int64_t stride, int64_t m, int64_t n, int32_t* nobs, int64_t* sum,
                      int64_t* sumsq, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_snp_stats, dim3((unsigned)ceil_div(m, 4)), dim3(256), 0, st, snp, stride, m, n, nobs, sum,
                       sumsq);
    GPE_HIP(hipGetLastError());
}

// Mean / inverse sd: mean-imputed (genotype.cpp:652-676) or completed data
// (genotype.cpp:629-650).
__global__ void k_std_params(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                             int ddof, int completed, double* mean, double* inv_sd) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const int64_t cnt = completed ? n : (int64_t)nobs[j];
    const double dn = (double)cnt;
    const double s = (double)sum[j];
    const double ss = (double)sumsq[j];
    double mu = 0.0, M2 = 0.0, inv = 1.0;
    if (cnt > 0) {
        mu = s / dn;
        M2 = ss - (s * s) / dn;
        if (M2 < 0.0 && M2 > -1e-12) M2 = 0.0;
    }
    const int64_t denom = cnt - (int64_t)ddof;
    if (denom > 0 && M2 > 0.0) {
        const double var = M2 / (double)denom;
        if (var > 0.0 && isfinite(var)) inv = 1.0 / sqrt(var);
    }
    mean[j] = mu;
    inv_sd[j] = inv;
}
void launch_std_params(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                       int ddof, int completed, double* mean, double* inv_sd, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_std_params, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, nobs, sum, sumsq, m, n,
                       ddof, completed, mean, inv_sd);
    GPE_HIP(hipGetLastError());
}

// Residual scale 1/sqrt(max(||x_j||^2/(N-1), eps)) with ||x_j||^2 from the
// integer moments (gwldcore.cpp:3220-3333, no covariates).
__global__ void k_resvar_inv(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* mean,
                             const double* inv_sd, int64_t r0, int64_t r1, int64_t n, int completed, double eps,
                             double* inv_out) {
    const int64_t j = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= r1) return;
    const double cnt = completed ? (double)n : (double)nobs[j];
    const double mu = mean[j];
    const double s = (double)sum[j];
    const double ss = (double)sumsq[j];
    double M2 = cnt > 0 ? ss - 2.0 * mu * s + mu * mu * cnt : 0.0;
    if (M2 < 0.0) M2 = 0.0;
    const double isd = inv_sd[j];
    const double sq = M2 * isd * isd;
    const double var = sq / (double)(n - 1);
    double safe;
    if (isnan(var)) safe = var;
    else safe = var > eps ? var : eps;
    inv_out[j - r0] = 1.0 / sqrt(safe);
}
void launch_resvar_inv(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* mean,
                       const double* inv_sd, int64_t r0, int64_t r1, int64_t n, int completed, double eps,
                       double* inv_out, hipStream_t st) {
    if (r1 <= r0) return;
    hipLaunchKernelGGL(k_resvar_inv, dim3((unsigned)ceil_div(r1 - r0, 256)), dim3(256), 0, st, nobs, sum, sumsq,
                       mean, inv_sd, r0, r1, n, completed, eps, inv_out);
    GPE_HIP(hipGetLastError());
}

// Deterministic HWE completion of missing calls (genotype.cpp:122-184, :576-595):
// per (SNP, selected row) a splitmix hash of (impute_seed, global SNP, global
// row) is compared with the HWE genotype frequencies of the observed calls.
__global__ void k_hwe_impute(uint8_t* snp, int64_t stride, int64_t m, int64_t n, const int32_t* nobs,
                             const int64_t* sum, uint64_t impute_seed, int64_t snp_offset, const int64_t* rows) {
    const int64_t words = stride / 4;
    const int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = blockIdx.y + (int64_t)blockIdx.z * 65535;
    if (w >= words || j >= m) return;
    uint32_t* pw = reinterpret_cast<uint32_t*>(snp + j * stride) + w;
    uint32_t x = *pw;
    const uint32_t lo = x & 0x55555555u, hi = (x >> 1) & 0x55555555u;
    uint32_t miss = lo & ~hi;
    if (!miss) return;
    const int64_t no = nobs[j];
    double p = 0.0;
    if (no > 0) {
        p = (double)sum[j] / (2.0 * (double)no);
        if (p < 0.0) p = 0.0;
        if (p > 1.0) p = 1.0;
    }
    const double q0 = (1.0 - p) * (1.0 - p);
    const double q01 = q0 + 2.0 * p * (1.0 - p);
    uint64_t base = mix64(impute_seed);
    base ^= mix64((uint64_t)(uint32_t)(int32_t)(snp_offset + j) + 0x243f6a8885a308d3ULL);
    while (miss) {
        const int bit = __ffs(miss) - 1;  // even bit position of the code's low bit
        miss &= miss - 1;
        const int t = bit >> 1;
        const int64_t smp = w * 16 + t;
        if (smp >= n) continue;
        const int64_t grow = rows ? rows[smp] : smp;
        uint32_t val;
        if (q0 >= 1.0) {
            val = 0;
        } else if (q01 <= 0.0) {
            val = 2;
        } else {
            uint64_t h = base ^ mix64((uint64_t)(uint32_t)(int32_t)grow + 0x9e3779b97f4a7c15ULL);
            const uint64_t r = mix64(h);
            const double u = (double)((r >> 11) & ((1ULL << 53) - 1)) * (1.0 / 9007199254740992.0);
            val = u < q0 ? 0u : (u < q01 ? 1u : 2u);
        }
        const uint32_t code = val == 0 ? 0u : (val == 1 ? 2u : 3u);
        x = (x & ~(3u << (2 * t))) | (code << (2 * t));
    }
    *pw = x;
}
void launch_hwe_impute(uint8_t* snp, int64_t stride, int64_t m, int64_t n, const int32_t* nobs,
                       const int64_t* sum, uint64_t impute_seed, int64_t snp_offset, const int64_t* rows,
                       hipStream_t st) {
    if (m <= 0) return;
    const int64_t words = stride / 4;
    dim3 grid((unsigned)ceil_div(words, 256), (unsigned)(m < 65535 ? m : 65535), (unsigned)ceil_div(m, 65535));
    hipLaunchKernelGGL(k_hwe_impute, grid, dim3(256), 0, st, snp, stride, m, n, nobs, sum, impute_seed, snp_offset,
                       rows);
    GPE_HIP(hipGetLastError());
}

// SNP-major -> sample-major transpose of 2-bit codes.  Tile: 64 SNPs x 256 samples.
__global__ __launch_bounds__(256) void k_transpose(const uint8_t* snp, int64_t snp_stride, uint8_t* smp,
                                                   int64_t smp_stride) {
    __shared__ uint8_t codes[64][257];
    const int tid = threadIdx.x;
    const int64_t n0 = (int64_t)blockIdx.x * 256;
    const int64_t j0 = (int64_t)blockIdx.y * 64;
    {
        const int row = tid >> 2;
        const int part = tid & 3;
        const uint4 v = *reinterpret_cast<const uint4*>(snp + (j0 + row) * snp_stride + n0 / 4 + part * 16);
        const uint32_t ws[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < 16; ++t) codes[row][part * 64 + q * 16 + t] = (uint8_t)((ws[q] >> (2 * t)) & 3u);
    }
    __syncthreads();
    uint32_t out[4] = {0, 0, 0, 0};
#pragma unroll
    for (int r = 0; r < 64; ++r) out[r >> 4] |= (uint32_t)codes[r][tid] << (2 * (r & 15));
    *reinterpret_cast<uint4*>(smp + (n0 + tid) * smp_stride + j0 / 4) = make_uint4(out[0], out[1], out[2], out[3]);
}
void launch_transpose(const uint8_t* snp, int64_t snp_stride, int64_t m, uint8_t* smp, int64_t smp_stride,
                      int64_t n, hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const int64_t gx = ceil_div(n, 256), gy = ceil_div(m, 64);
    for (int64_t y0 = 0; y0 < gy; y0 += 65535) {
        const int64_t yy = (gy - y0) < 65535 ? (gy - y0) : 65535;
        hipLaunchKernelGGL(k_transpose, dim3((unsigned)gx, (unsigned)yy), dim3(256), 0, st,
                           snp + y0 * 64 * snp_stride, snp_stride, smp + y0 * 16, smp_stride);
        GPE_HIP(hipGetLastError());
    }
}

// Missing calls per sample row of the sample-major layout.
__global__ void k_miss_count(const uint8_t* smp, int64_t stride, int64_t n, int32_t* counts) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(smp + r * stride);
    int64_t c = 0;
    for (int64_t w = lane; w < stride / 4; w += 64) {
        const uint32_t x = row[w];
        c += __popc((x & 0x55555555u) & ~((x >> 1) & 0x55555555u));
    }
    c = wave_sum_i64(c);
    if (lane == 0) counts[r] = (int32_t)c;
}
void launch_miss_count(const uint8_t* smp, int64_t stride, int64_t n, int32_t* counts, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_miss_count, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st, smp, stride, n, counts);
    GPE_HIP(hipGetLastError());
}

// Sorted indices of the missing calls of each packed row: SNP indices for
// sample-major rows, sample indices for SNP-major rows.
__global__ void k_miss_fill(const uint8_t* smp, int64_t stride, int64_t n, const int64_t* ptr, int32_t* idx) {
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (r >= n) return;
    const uint32_t* row = reinterpret_cast<const uint32_t*>(smp + r * stride);
    int64_t pos = ptr[r];
    const int64_t words = stride / 4;
    for (int64_t w0 = 0; w0 < words; w0 += 64) {
        const int64_t w = w0 + lane;
        uint32_t mm = 0;
        if (w < words) {
            const uint32_t x = row[w];
            mm = (x & 0x55555555u) & ~((x >> 1) & 0x55555555u);
        }
        const int cnt = __popc(mm);
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        int64_t p = pos + incl - cnt;
        while (mm) {
            const int bit = __ffs(mm) - 1;
            mm &= mm - 1;
            idx[p++] = (int32_t)(w * 16 + (bit >> 1));
        }
        pos += __shfl(incl, 63, 64);
    }
}
void launch_miss_fill(const uint8_t* smp, int64_t stride, int64_t n, const int64_t* ptr, int32_t* idx,
                      hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_miss_fill, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st, smp, stride, n, ptr, idx);
    GPE_HIP(hipGetLastError());
}

// Exclusive scan of n counts in one workgroup: each thread sums a contiguous
// run, the 1024 run totals are scanned in LDS, then each run is written out.
__global__ __launch_bounds__(1024) void k_scan_counts(const int32_t* __restrict__ counts, int64_t n,
                                                      int64_t* __restrict__ ptr) {
    __shared__ int64_t part[1024];
    const int t = threadIdx.x;
    const int64_t per = (n + 1023) / 1024;
    const int64_t a = t * per < n ? t * per : n;
    const int64_t b = a + per < n ? a + per : n;
    int64_t s = 0;
    for (int64_t i = a; i < b; ++i) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {
        const int64_t v = t >= o ? part[t - o] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t acc = part[t] - s;
    for (int64_t i = a; i < b; ++i) {
        ptr[i] = acc;
        acc += counts[i];
    }
    if (t == 1023) ptr[n] = part[1023];
}
void launch_scan_counts(const int32_t* counts, int64_t n, int64_t* ptr, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_counts, dim3(1), dim3(1024), 0, st, counts, n, ptr);
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Probe streams on the GPU: std::mt19937_64 seeded with make_seed(root, block,
// v_start) per SNP block (gwldcore.cpp:445-458, :1691-1716).  One wave per
// block stream; the 312-word twist runs lane-parallel in three dependency
// phases, and output t of the stream is probe t / L of SNP block_start + t % L.
// ---------------------------------------------------------------------------
constexpr int kMtN = 312;
constexpr int kMtM = 156;
constexpr int kMtStage = 65536;  // bytes of a probe tile staged in LDS
constexpr uint64_t kMtA = 0xB5026F5AA96619E9ULL;
constexpr uint64_t kMtUM = 0xFFFFFFFF80000000ULL;
constexpr uint64_t kMtLM = 0x7FFFFFFFULL;

__device__ __forceinline__ uint64_t mt_twist_one(uint64_t a, uint64_t b, uint64_t c) {
    const uint64_t x = (a & kMtUM) | (b & kMtLM);
    return c ^ (x >> 1) ^ ((x & 1ULL) ? kMtA : 0ULL);
}

// One wave per (SNP block, probe tile): the stream of the reference's V-tile
// [v_start + t*sub, + sub) of that block, written to probe columns t*sub.. .
__global__ __launch_bounds__(64) void k_mt_rademacher(uint64_t root, const int64_t* blk_starts, int n_blocks,
                                                     int64_t blk_end_all, int v_start_arg, const int32_t* v_start_dev,
                                                     int v_count_all, int sub, int8_t* z_all, int64_t ldz) {
    __shared__ uint64_t mt[kMtN];
    // the block's probe tile, row-major (L, v_count), when it fits: the draws fill it
    // column by column and it is written out as whole rows (contiguous stores)
    // instead of one byte per draw at a stride of ldz
    __shared__ __attribute__((aligned(16))) int8_t stage[kMtStage];
    const int b = blockIdx.x;
    const int lane = threadIdx.x;
    if (b >= n_blocks) return;
    const int v_start = v_start_dev ? v_start_dev[0] : v_start_arg;
    const int v0 = (int)blockIdx.y * sub;
    if (v0 >= v_count_all) return;
    const int v_count = v_count_all - v0 < sub ? v_count_all - v0 : sub;
    int8_t* z = z_all + v0;
    const int64_t s = blk_starts[b];
    const int64_t e = (b + 1 < n_blocks) ? blk_starts[b + 1] : blk_end_all;
    const int64_t L = e - s;
    if (lane == 0) {
        uint64_t seed = 0x1234abcdULL;
        seed ^= mix64(root);
        seed ^= mix64((uint64_t)(int64_t)(int32_t)s + 0x9e37ULL);
        seed ^= mix64((uint64_t)(int64_t)(v_start + v0) + 0x85ebULL);
        mt[0] = seed;
        for (int i = 1; i < kMtN; ++i) mt[i] = 6364136223846793005ULL * (mt[i - 1] ^ (mt[i - 1] >> 62)) + (uint64_t)i;
    }
    __syncthreads();
    const int64_t total = L * (int64_t)v_count;
    const bool staged = total <= kMtStage;
    // output t = v L + r of this lane's slots i = lane + 64 q, advanced by kMtN per
    // twist without a division
    int32_t ov[5], orow[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) {
        const int i = lane + 64 * q;
        ov[q] = (int32_t)(i / L);
        orow[q] = (int32_t)(i % L);
    }
    // the state in registers: lane l holds mt[l + 64 q] in m[q] (q < 5, index < 312);
    // a twist exchanges the neighbours it needs across lanes instead of through LDS
    uint64_t m[5];
#pragma unroll
    for (int q = 0; q < 5; ++q) m[q] = lane + 64 * q < kMtN ? mt[lane + 64 * q] : 0ULL;
    auto shfl64 = [](uint64_t v, int src) -> uint64_t {
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
        return ((uint64_t)hi << 32) | lo;
    };
    auto lane64 = [](uint64_t v, int src) -> uint64_t {
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, src);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), src);
        return ((uint64_t)hi << 32) | lo;
    };
    const int up1 = (lane + 1) & 63;
    const int up28 = (lane + 28) & 63;  // i + 156 = (lane + 28) + 64 (q + 2 or q + 3)
    const int up36 = (lane + 36) & 63;  // i - 156 = (lane + 36) + 64 (q - 3 or q - 2)
    for (int64_t base = 0; base < total; base += kMtN) {
        // old mt[i + 1] of every slot
        uint64_t nb[5];
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const uint64_t x = shfl64(m[q], up1);
            const uint64_t w = q < 4 ? lane64(m[q + 1], 0) : 0ULL;
            nb[q] = lane == 63 ? w : x;
        }
        // i < 156 (slots 0, 1; slot 2 lanes < 28) from old mt[i + 156]; every exchange
        // runs with all lanes active (a lane-permute reads nothing from an inactive lane)
        const uint64_t g2 = shfl64(m[2], up28), g3 = shfl64(m[3], up28), g4 = shfl64(m[4], up28);
        const uint64_t f0 = lane < 36 ? g2 : g3;
        const uint64_t f1 = lane < 36 ? g3 : g4;
        const uint64_t f2 = g4;
        const uint64_t a0 = mt_twist_one(m[0], nb[0], f0);
        const uint64_t a1 = mt_twist_one(m[1], nb[1], f1);
        const uint64_t a2 = mt_twist_one(m[2], nb[2], f2);
        // 156 <= i < 311 (slot 2 lanes >= 28, slot 3, slot 4 lanes < 55) from new mt[i - 156]
        const uint64_t s0 = shfl64(a0, up36), s1 = shfl64(a1, up36), s2 = shfl64(a2, up36);
        const uint64_t b2 = mt_twist_one(m[2], nb[2], s0);
        const uint64_t b3 = mt_twist_one(m[3], nb[3], lane < 28 ? s0 : s1);
        const uint64_t b4 = mt_twist_one(m[4], nb[4], lane < 28 ? s1 : s2);
        // i = 311 from old mt[311], new mt[0] and new mt[155]
        const uint64_t n311 = mt_twist_one(lane64(m[4], 55), lane64(a0, 0), lane64(a2, 27));
        m[0] = a0;
        m[1] = a1;
        m[2] = lane < 28 ? a2 : b2;
        m[3] = b3;
        m[4] = lane < 55 ? b4 : (lane == 55 ? n311 : 0ULL);
#pragma unroll
        for (int q = 0; q < 5; ++q) {
            const int i = lane + 64 * q;
            if (i < kMtN && ov[q] < v_count) {
                uint64_t y = m[q];
                y ^= (y >> 29) & 0x5555555555555555ULL;
                y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
                y ^= (y << 37) & 0xFFF7EEE000000000ULL;
                y ^= (y >> 43);
                const int8_t val = (y & 1ULL) ? (int8_t)1 : (int8_t)-1;
                if (staged)
                    stage[orow[q] * v_count + ov[q]] = val;
                else
                    z[(s + orow[q]) * ldz + ov[q]] = val;
            }
            orow[q] += kMtN;
            while (orow[q] >= L) {
                orow[q] -= (int32_t)L;
                ++ov[q];
            }
        }
    }
    __syncthreads();
    if (staged) {
        // whole dwords when the rows allow it, else bytes (consecutive lanes, consecutive bytes)
        if (v_count % 4 == 0 && ldz % 4 == 0 && (reinterpret_cast<uintptr_t>(z) & 3) == 0) {
            const int dpr = v_count / 4;
            for (int64_t d = lane; d < total / 4; d += 64) {
                const int64_t r = d / dpr;
                const int c = (int)(d - r * dpr);
                *reinterpret_cast<uint32_t*>(z + (s + r) * ldz + 4 * c) = reinterpret_cast<const uint32_t*>(stage)[d];
            }
        } else {
            for (int64_t d = lane; d < total; d += 64) z[(s + d / v_count) * ldz + d % v_count] = stage[d];
        }
    }
}
void launch_mt_rademacher(uint64_t root, const int64_t* blk_starts, int n_blocks, int64_t blk_end_all,
                          int v_start, int v_count, int8_t* z, int64_t ldz, hipStream_t st, int sub,
                          const int32_t* v_start_dev) {
    if (n_blocks <= 0 || v_count <= 0) return;
    if (sub <= 0 || sub > v_count) sub = v_count;
    const int n_sub = (int)ceil_div(v_count, sub);
    hipLaunchKernelGGL(k_mt_rademacher, dim3((unsigned)n_blocks, (unsigned)n_sub), dim3(64), 0, st, root, blk_starts,
                       n_blocks, blk_end_all, v_start, v_start_dev, v_count, sub, z, ldz);
    GPE_HIP(hipGetLastError());
}

__global__ void k_add_i32(int32_t* v, int32_t inc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) v[0] += inc;
}

void launch_add_i32(int32_t* v, int32_t inc, hipStream_t st) {
    hipLaunchKernelGGL(k_add_i32, dim3(1), dim3(64), 0, st, v, inc);
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Quantization of real coefficients into int8 limbs
// ---------------------------------------------------------------------------
__device__ __forceinline__ double p1_coef(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64,
                                          int64_t ldz, int64_t i, int k, int v) {
    const double zz = z_i8 ? (double)z_i8[i * ldz + v] : z_f64[i * ldz + v];
    return w[i * ldw + k] * zz;
}

// column max |coef| over SNPs [k0,k1); column c = bin * v_pad + v.  int8 probes
// hold -1, 0 or +1, so the bound max|w_ik| is used for every column of bin k:
// all columns of a bin then share one shift.
__global__ void k_coef_p1_colmax(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                                 int64_t k0, int64_t k1, int v_count, int v_pad, uint64_t* colmax_bits) {
    const int c = blockIdx.y;
    const int k = c / v_pad, v = c % v_pad;
    if (v >= v_count) return;
    double mx = 0.0;
    for (int64_t i = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k1; i += (int64_t)gridDim.x * blockDim.x)
        mx = fmax(mx, z_i8 ? fabs(w[i * ldw + k]) : fabs(p1_coef(w, ldw, z_i8, z_f64, ldz, i, k, v)));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0 && mx > 0.0)
        atomicMax(reinterpret_cast<unsigned long long*>(colmax_bits + c), (unsigned long long)__double_as_longlong(mx));
}
// int8 probes: every column of bin k gets max_i |w_ik| (see k_coef_p1_colmax); it
// is reduced once per bin into column k * v_pad, then copied to the bin's columns
__global__ void k_coef_p1_binmax(const double* w, int64_t ldw, int64_t k0, int64_t k1, int v_pad,
                                 uint64_t* colmax_bits) {
    const int k = blockIdx.y;
    double mx = 0.0;
    for (int64_t i = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k1; i += (int64_t)gridDim.x * blockDim.x)
        mx = fmax(mx, fabs(w[i * ldw + k]));
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    if ((threadIdx.x & 63) == 0 && mx > 0.0)
        atomicMax(reinterpret_cast<unsigned long long*>(colmax_bits + (int64_t)k * v_pad),
                  (unsigned long long)__double_as_longlong(mx));
}
__global__ void k_colmax_bin_copy(int n_bins, int v_count, int v_pad, uint64_t* colmax_bits) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n_bins * v_pad) return;
    const int k = c / v_pad, v = c % v_pad;
    if (v >= 1 && v < v_count) colmax_bits[c] = colmax_bits[(int64_t)k * v_pad];
}
void launch_coef_p1_colmax(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                           int64_t k0, int64_t k1, int n_bins, int v_count, int v_pad, uint64_t* colmax_bits,
                           hipStream_t st) {
    const int64_t L = k1 - k0;
    if (L <= 0) return;
    int64_t gx = ceil_div(L, 256);
    if (gx > 64) gx = 64;
    if (z_i8) {
        hipLaunchKernelGGL(k_coef_p1_binmax, dim3((unsigned)gx, (unsigned)n_bins), dim3(256), 0, st, w, ldw, k0, k1,
                           v_pad, colmax_bits);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_colmax_bin_copy, dim3((unsigned)ceil_div(n_bins * v_pad, 256)), dim3(256), 0, st, n_bins,
                           v_count, v_pad, colmax_bits);
        GPE_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_coef_p1_colmax, dim3((unsigned)gx, (unsigned)(n_bins * v_pad)), dim3(256), 0, st, w, ldw,
                       z_i8, z_f64, ldz, k0, k1, v_count, v_pad, colmax_bits);
    GPE_HIP(hipGetLastError());
}

// shift_c = P - exponent(max|x_c|) so that |x * 2^shift| <= 2^P
__global__ void k_shift_from_colmax(const uint64_t* colmax_bits, int n, int precision_bits, int32_t* shift,
                                    double* col_scale) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double mx = __longlong_as_double((long long)colmax_bits[c]);
    int sh = 0;
    if (mx > 0.0) {
        int e = 0;
        frexp(mx, &e);  // mx < 2^e
        sh = precision_bits - e;
    }
    shift[c] = sh;
    // a column holding a non-finite value (marked NaN by k_colmax_f64) gets a NaN
    // scale: every product with it is NaN, as a NaN in a dense FP64 GEMM operand
    // makes every dot product with that column NaN
    col_scale[c] = isnan(mx) ? __builtin_nan("") : ldexp(1.0, -sh);
}
void launch_shift_from_colmax(const uint64_t* colmax_bits, int n, int precision_bits, int32_t* shift,
                              double* col_scale, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_shift_from_colmax, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, st, colmax_bits, n,
                       precision_bits, shift, col_scale);
    GPE_HIP(hipGetLastError());
}

// Swizzled coefficient limbs of phase 1: thread (group g of 16 consecutive SNPs
// of the chunk range, column c) quantizes q = round(w_ik z_iv 2^shift_c) for the
// 16 SNPs (0 outside [k0, k1) or past v_count), places each limb's 16 digits at
// the transposed byte positions of one 16-byte operand unit and stores the unit
// whole; consecutive threads are consecutive columns (coalesced probe reads and
// unit stores).
template <int NL>
__global__ __launch_bounds__(256) void k_coef_p1_quant16(const double* w, int64_t ldw, const int8_t* z_i8,
                                                         const double* z_f64, int64_t ldz, int64_t k0, int64_t k1,
                                                         int v_count, int v_pad, int ncols, const int32_t* shift,
                                                         int ctc, int64_t coef_ctiles, int64_t kc0, int8_t* coef) {
    const int64_t grp = blockIdx.x;
    const int c = (int)blockIdx.y * 256 + (int)threadIdx.x;
    if (c >= ncols) return;
    const int k = c / v_pad, v = c % v_pad;
    const int sh = shift[c];
    uint32_t unit[NL][4];
#pragma unroll
    for (int l = 0; l < NL; ++l) unit[l][0] = unit[l][1] = unit[l][2] = unit[l][3] = 0u;
    const int64_t i0 = kc0 * kKChunk + grp * 16;
#pragma unroll 4
    for (int j = 0; j < 16; ++j) {
        const int64_t i = i0 + j;
        int64_t q = 0;
        if (v < v_count && i >= k0 && i < k1) q = __double2ll_rn(ldexp(p1_coef(w, ldw, z_i8, z_f64, ldz, i, k, v), sh));
        int8_t d[NL];
        balanced_digits<NL>(q, d);
        const int b = ((j & 3) << 2) | (j >> 2);  // the unpacked operand's byte order (swz_offset)
#pragma unroll
        for (int l = 0; l < NL; ++l) unit[l][b >> 2] |= (uint32_t)(uint8_t)d[l] << (8 * (b & 3));
    }
    const int64_t kl = grp / (kKChunk / 16);
    const int kk = (int)(grp % (kKChunk / 16)) * 16;
    const int ct = c >> 5, lr = c & 31;
    const int cgp = ct / ctc, ctcl = ct % ctc;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const int64_t tg = (int64_t)cgp * (NL * ctc) + l * ctc + ctcl;
        *reinterpret_cast<v4u*>(coef + swz_offset(kl, coef_ctiles, tg, kk, lr)) =
            v4u{unit[l][0], unit[l][1], unit[l][2], unit[l][3]};
    }
}

// Cst_c = sum_i mean_i * deq(coef_ic) in two deterministic passes: blockIdx.y
// = p sums its SNP slice into part[c * P + p], then k_coef_p1_cst_final adds
// the P partials of each column in a fixed order.
__global__ void k_coef_p1_cst_part(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                                   int64_t k0, int64_t k1, int v_count, int v_pad, const int32_t* shift,
                                   const double* mean, int64_t slice, double* part) {
    __shared__ double red[4];
    const int c = blockIdx.x;
    const int P = gridDim.y;
    const int k = c / v_pad, v = c % v_pad;
    const int64_t i0 = k0 + (int64_t)blockIdx.y * slice;
    const int64_t i1 = i0 + slice < k1 ? i0 + slice : k1;
    double acc = 0.0;
    if (v < v_count) {
        const int sh = shift[c];
        for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const double q = (double)__double2ll_rn(ldexp(p1_coef(w, ldw, z_i8, z_f64, ldz, i, k, v), sh));
            acc += mean[i] * q;
        }
    }
    acc = wave_sum_f64(acc);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) part[(int64_t)c * P + blockIdx.y] = (red[0] + red[1]) + (red[2] + red[3]);
}
__global__ void k_coef_p1_cst_final(const double* part, int P, int v_count, int v_pad, const int32_t* shift,
                                    double* cst) {
    const int c = blockIdx.x;
    const int v = c % v_pad;
    double acc = 0.0;
    for (int p = threadIdx.x; p < P; p += 64) acc += part[(int64_t)c * P + p];
    acc = wave_sum_f64(acc);
    if (threadIdx.x == 0) cst[c] = v < v_count ? ldexp(acc, -shift[c]) : 0.0;
}

void launch_coef_p1_quant(const double* w, int64_t ldw, const int8_t* z_i8, const double* z_f64, int64_t ldz,
                          int64_t k0, int64_t k1, int n_bins, int v_count, int v_pad, const int32_t* shift, int nl,
                          int ctc, int64_t coef_ctiles, int64_t kc0, int8_t* coef, const double* mean, double* cst,
                          double* cst_part, hipStream_t st) {
    const int64_t kc1 = ceil_div(k1, kKChunk);
    const int64_t kcount = kc1 - kc0;
    const int ncols = n_bins * v_pad;
    if (kcount > 0) {
        // one thread per (16 consecutive SNPs, column): the 16 values of one 16-byte
        // operand unit per limb, written with one vector store each
        const dim3 grid((unsigned)(kcount * (kKChunk / 16)), (unsigned)ceil_div(ncols, 256));
        if (nl == 5)
            hipLaunchKernelGGL(k_coef_p1_quant16<5>, grid, dim3(256), 0, st, w, ldw, z_i8, z_f64, ldz, k0, k1,
                               v_count, v_pad, ncols, shift, ctc, coef_ctiles, kc0, coef);
        else if (nl == 6)
            hipLaunchKernelGGL(k_coef_p1_quant16<6>, grid, dim3(256), 0, st, w, ldw, z_i8, z_f64, ldz, k0, k1,
                               v_count, v_pad, ncols, shift, ctc, coef_ctiles, kc0, coef);
        else
            throw Error("phase-1 coefficient limbs must be 5 or 6");
        GPE_HIP(hipGetLastError());
    }
    if (cst) {
        const int64_t slice = std::max<int64_t>(4096, ceil_div(k1 - k0, kCstParts));
        const int P = (int)std::max<int64_t>(1, ceil_div(k1 - k0, slice));
        hipLaunchKernelGGL(k_coef_p1_cst_part, dim3((unsigned)ncols, (unsigned)P), dim3(256), 0, st, w, ldw, z_i8,
                           z_f64, ldz, k0, k1, v_count, v_pad, shift, mean, slice, cst_part);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_coef_p1_cst_final, dim3((unsigned)ncols), dim3(64), 0, st, cst_part, P, v_count, v_pad,
                           shift, cst);
        GPE_HIP(hipGetLastError());
    }
}

// Phase-1 finalize: Xz[:, bin*V+v] (+)= Y*2^-shift - Cst + sum_{missing i} mean_i * coef_i.
// One wave per sample, lanes over columns.
__global__ void k_p1_finalize(const int64_t* y, int64_t ldy, int64_t n, int n_bins, int v_count, int v_pad,
                              const double* col_scale, const double* cst, const int64_t* miss_ptr,
                              const int32_t* miss_idx, int64_t k0, int64_t k1, const double* w, int64_t ldw,
                              const int8_t* z_i8, const double* z_f64, int64_t ldz, const int32_t* shift,
                              const double* mean, double* xz, int64_t ldxz, const uint8_t* col_mode) {
    const int64_t smp = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (smp >= n) return;  // wave-uniform
    int64_t mb = 0, me = 0;
    if (miss_ptr) {
        // entries of this sample's sorted list inside [k0, k1)
        const int64_t b0 = miss_ptr[smp], b1 = miss_ptr[smp + 1];
        int64_t lo = b0, hi = b1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (miss_idx[mid] < k0) lo = mid + 1;
            else hi = mid;
        }
        mb = lo;
        hi = b1;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (miss_idx[mid] < k1) lo = mid + 1;
            else hi = mid;
        }
        me = lo;
    }
    const int ncols = n_bins * v_pad;
    for (int c0 = 0; c0 < ncols; c0 += 64) {
        const int c = c0 + lane;
        const int k = c / v_pad, v = c % v_pad;
        const bool valid = c < ncols && v < v_count;
        double val = valid ? (double)y[(int64_t)c * ldy + smp] * col_scale[c] - cst[c] : 0.0;
        if (miss_ptr) {
            const int sh = valid ? shift[c] : 0;
            const double sc = valid ? col_scale[c] : 0.0;
            double acc = 0.0;
            // indices are loaded 64 at a time and broadcast lane by lane, so the
            // per-entry loads are independent and pipeline
            for (int64_t p0 = mb; p0 < me; p0 += 64) {
                const int cnt = (me - p0) < 64 ? (int)(me - p0) : 64;
                const int32_t myidx = (p0 + lane < me) ? miss_idx[p0 + lane] : 0;
#pragma unroll 8
                for (int j = 0; j < cnt; ++j) {
                    const int64_t i = __builtin_amdgcn_readlane(myidx, j);
                    if (valid) {
                        const double q =
                            (double)__double2ll_rn(ldexp(p1_coef(w, ldw, z_i8, z_f64, ldz, i, k, v), sh));
                        acc += mean[i] * (q * sc);
                    }
                }
            }
            val += acc;
        }
        if (valid) {
            const int mode = col_mode ? col_mode[k] : 2;
            double* dst = xz + (int64_t)(k * v_count + v) * ldxz + smp;
            if (mode == 1) *dst = val;
            else *dst += val;
        }
    }
}
__global__ void k_p1_combine2(const int64_t* __restrict__ y, const int64_t* __restrict__ y2, int64_t ldy, int64_t n,
                              int n_bins, int v_count, int v_pad, const double* __restrict__ cs,
                              const double* __restrict__ cst, const double* __restrict__ cs2, double* __restrict__ xz,
                              int64_t ldxz, const uint8_t* __restrict__ col_mode) {
    const int64_t smp = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    const int k = c / v_pad, v = c % v_pad;
    if (smp >= n || v >= v_count) return;
    const double val = (double)y[(int64_t)c * ldy + smp] * cs[c] - cst[c] + (double)y2[(int64_t)c * ldy + smp] * cs2[c];
    const int mode = col_mode ? col_mode[k] : 2;
    double* dst = xz + (int64_t)(k * v_count + v) * ldxz + smp;
    if (mode == 1) *dst = val;
    else *dst += val;
}

void launch_p1_combine2(const int64_t* y, const int64_t* y2, int64_t ldy, int64_t n, int n_bins, int v_count,
                        int v_pad, const double* cs, const double* cst, const double* cs2, double* xz, int64_t ldxz,
                        const uint8_t* col_mode, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_p1_combine2, dim3((unsigned)ceil_div(n, 256), (unsigned)(n_bins * v_pad)), dim3(256), 0, st,
                       y, y2, ldy, n, n_bins, v_count, v_pad, cs, cst, cs2, xz, ldxz, col_mode);
    GPE_HIP(hipGetLastError());
}

__global__ void k_scale_rows(const double* __restrict__ w, const double* __restrict__ mean, int64_t s, int64_t e,
                             int n_bins, double* __restrict__ out) {
    const int64_t idx = s * n_bins + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= e * n_bins) return;
    out[idx] = mean[idx / n_bins] * w[idx];
}

void launch_scale_rows(const double* w, const double* mean, int64_t s, int64_t e, int n_bins, double* out,
                       hipStream_t st) {
    const int64_t total = (e - s) * (int64_t)n_bins;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_scale_rows, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, w, mean, s, e, n_bins,
                       out);
    GPE_HIP(hipGetLastError());
}

void launch_p1_finalize(const int64_t* y, int64_t ldy, int64_t n, int n_bins, int v_count, int v_pad,
                        const double* col_scale, const double* cst, const int64_t* miss_ptr,
                        const int32_t* miss_idx, int64_t k0, int64_t k1, const double* w, int64_t ldw,
                        const int8_t* z_i8, const double* z_f64, int64_t ldz, const int32_t* shift,
                        const double* mean, double* xz, int64_t ldxz, const uint8_t* col_mode, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_p1_finalize, dim3((unsigned)ceil_div(n, 4)), dim3(256), 0, st, y, ldy, n, n_bins, v_count,
                       v_pad, col_scale, cst, miss_ptr, miss_idx, k0, k1, w, ldw, z_i8, z_f64, ldz, shift, mean, xz,
                       ldxz, col_mode);
    GPE_HIP(hipGetLastError());
}

// First position p in [lo, hi) of the sorted idx with idx[p] >= t (hi if none),
// searched by the whole wave: each step the 64 lanes probe 64 evenly spaced
// positions and a ballot of idx < t (a prefix of the lanes, idx being sorted)
// narrows the interval about 65-fold, so a sample's list of thousands of
// entries takes two or three dependent loads instead of a dozen.  Invariant:
// the answer lies in [lo, hi]; positions below lo hold idx < t and position
// hi, when below the original end, holds idx >= t.  lo, hi, t wave-uniform.
__device__ __forceinline__ int64_t wave_lower_bound(const int32_t* __restrict__ idx, int64_t lo, int64_t hi,
                                                    int64_t t, int lane) {
    while (hi - lo > 64) {
        const int64_t len = hi - lo;
        const int64_t pos = lo + (len * (lane + 1)) / 65;  // in [lo, hi), increasing with the lane
        const int cnt = __popcll(__ballot(idx[pos] < t));
        const int64_t nlo = cnt > 0 ? lo + (len * cnt) / 65 + 1 : lo;
        const int64_t nhi = cnt < 64 ? lo + (len * (cnt + 1)) / 65 : hi;
        lo = nlo;
        hi = nhi;
    }
    const bool less = lo + lane < hi && idx[lo + lane] < t;
    return lo + __popcll(__ballot(less));
}

// Missing-call records for phase 1 with int8 probes (entries -1, 0, +1): for
// SNP i of [k0, k1) and probe word wd (columns 64wd .. 64wd+63) one record of
// rec_bytes = round_up(16 + 8 n_bins, 16) bytes at byte ((i - k0) vw + wd) rec_bytes:
//   uint32 code words t = 0..3, column 64wd + 16t + b in bits 2b, 2b+1 as a
//   signed 2-bit field (+1 -> 01, -1 -> 11, 0 -> 00: one signed bit-field
//   extract gives z), then double a_k for every bin k,
// a_k = mean_i * (round(w_ik 2^shift_k) * 2^-shift_k).
// Rounding to nearest even commutes with negation, so the quantized
// coefficient of column (k, v) is z_iv round(w_ik 2^shift_k) and
// mean_i * deq(coef) is z_iv a_k bit for bit.
__host__ __device__ constexpr int64_t miss_rec_bytes(int n_bins) { return ((16 + 8 * (int64_t)n_bins + 15) / 16) * 16; }

// One thread per (SNP, 64-probe word): the word's 64 probe bytes in four
// 16-byte loads (the probe rows are 32-byte aligned), code words of the first
// v_count, then the bins' a_k.
__global__ void k_p1_miss_rec(const double* w, int64_t ldw, const int8_t* z, int64_t ldz, int64_t k0, int64_t k1,
                              int n_bins, int v_count, int v_pad, const int32_t* shift, const double* col_scale,
                              const double* mean, uint8_t* rec) {
    const int vw = (v_pad + 63) / 64;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (k1 - k0) * vw) return;
    const int64_t i = k0 + t / vw;
    const int wd = (int)(t % vw);
    const int64_t rb = miss_rec_bytes(n_bins);
    uint32_t code[4] = {0u, 0u, 0u, 0u};
    const int nb = v_count - wd * 64 < 64 ? v_count - wd * 64 : 64;  // probe columns of this word
    const int8_t* zr = z + i * ldz + wd * 64;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q * 16 >= nb) break;
        const v4u d = *reinterpret_cast<const v4u*>(zr + 16 * q);
#pragma unroll
        for (int b = 0; b < 16; ++b) {
            const int8_t zz = (int8_t)((d[b >> 2] >> (8 * (b & 3))) & 0xffu);
            if (q * 16 + b < nb) {
                if (zz > 0) code[q] |= 1u << (2 * b);
                else if (zz < 0) code[q] |= 3u << (2 * b);
            }
        }
    }
    uint8_t* r = rec + ((i - k0) * vw + wd) * rb;
    *reinterpret_cast<v4u*>(r) = v4u{code[0], code[1], code[2], code[3]};
    double* a = reinterpret_cast<double*>(r + 16);
    for (int k = 0; k < n_bins; ++k) {
        const int c = k * v_pad;
        const double q = (double)__double2ll_rn(ldexp(w[i * ldw + k], shift[c]));
        a[k] = mean[i] * (q * col_scale[c]);
    }
    for (int64_t b = 16 + 8 * (int64_t)n_bins; b < rb; ++b) r[b] = 0;
}
void launch_p1_miss_rec(const double* w, int64_t ldw, const int8_t* z, int64_t ldz, int64_t k0, int64_t k1,
                        int n_bins, int v_count, int v_pad, const int32_t* shift, const double* col_scale,
                        const double* mean, void* rec, hipStream_t st) {
    if (k1 <= k0) return;
    if (ldz % 16 != 0) throw Error("probe rows must be 16-byte aligned");
    const int64_t items = (k1 - k0) * ceil_div(v_pad, 64);
    hipLaunchKernelGGL(k_p1_miss_rec, dim3((unsigned)ceil_div(items, 256)), dim3(256), 0, st, w, ldw, z, ldz, k0,
                       k1, n_bins, v_count, v_pad, shift, col_scale, mean, static_cast<uint8_t*>(rec));
    GPE_HIP(hipGetLastError());
}
int64_t p1_miss_rec_bytes(int64_t n_snps, int n_bins, int v_pad) {
    return n_snps * ceil_div(v_pad, 64) * miss_rec_bytes(n_bins);
}

// Phase-1 finalize for int8 probes with missing calls: as k_p1_finalize, with
// the missing-call term sum_{missing i} z_iv a_ik taken from the records of
// k_p1_miss_rec for KB bins [kb0, kb0 + KB) at once.  One wave per sample, lanes
// over 64 probe columns per pass.  The sample's entries are handled 64 at a
// time: lane j loads entry j's record (the next group's loads are in flight
// while the current group is summed) and stages it in the wave's LDS slot;
// every lane then reads each record back (its column's code word once, the KB
// shared a_k) and adds z * a_k with one FMA per bin (z in {-1, 0, +1}, exact).
constexpr int kMissGroup = 64;

template <int KB>
__global__ __launch_bounds__(256) void k_p1_finalize_z8(
    const int64_t* __restrict__ y, int64_t ldy, int64_t n, int n_bins, int kb0, int v_count, int v_pad,
    const double* __restrict__ col_scale, const double* __restrict__ cst, const int64_t* __restrict__ miss_ptr,
    const int32_t* __restrict__ miss_idx, int64_t k0, int64_t ra, int64_t rb, const uint8_t* __restrict__ rec,
    double* __restrict__ xz, int64_t ldxz, const uint8_t* __restrict__ col_mode) {
    constexpr int kStage = ((16 + 8 * KB + 15) / 16) * 16;  // staged bytes per entry
    constexpr int kPieces = kStage / 16;
    __shared__ __attribute__((aligned(16))) uint8_t stage[4][kMissGroup * kStage];
    // this launch adds the missing calls at SNPs [ra, rb); the one with ra == k0
    // also applies Y 2^-shift - Cst and the bins' overwrite/accumulate modes
    const bool first = ra == k0;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t smp = (int64_t)blockIdx.x * 4 + wave;
    const int lane = threadIdx.x & 63;
    if (smp >= n) return;
    const int64_t b0 = miss_ptr[smp], b1 = miss_ptr[smp + 1];
    const int64_t mb = wave_lower_bound(miss_idx, b0, b1, ra, lane);
    const int64_t me = wave_lower_bound(miss_idx, mb, b1, rb, lane);
    const int vw = (v_pad + 63) / 64;
    const int64_t rbytes = miss_rec_bytes(n_bins);
    const int cword = lane >> 4;                         // code word of this lane's column
    const uint32_t cshift = 30u - 2u * (uint32_t)(lane & 15);  // its field moved to bits 30-31
    uint8_t* st = stage[wave];
    for (int wd = 0; wd < vw; ++wd) {
        const int v = wd * 64 + lane;
        const bool valid = v < v_count;
        // byte offset of this pass's part of SNP i's record: header + a[kb0 .. kb0+KB)
        const uint8_t* rbase = rec + (int64_t)wd * rbytes;
        double acc[KB];
#pragma unroll
        for (int k = 0; k < KB; ++k) acc[k] = 0.0;
        // a lane past the end reads the record of SNP k0 (in bounds) and stages nothing used
        auto entry_ptr = [&](int64_t p) -> const uint8_t* {
            const int32_t i = (p < me) ? miss_idx[p] : (int32_t)k0;
            return rbase + ((int64_t)i - k0) * vw * rbytes;
        };
        v4u cur[kPieces];
        {
            const uint8_t* e = entry_ptr(mb + lane);
            cur[0] = *reinterpret_cast<const v4u*>(e);
#pragma unroll
            for (int q = 1; q < kPieces; ++q) {
                // a values kb0.. of this pass; bytes past the record's bins are never read back
                const int64_t off = 16 + 8 * (int64_t)kb0 + 16 * (q - 1);
                cur[q] = off + 16 <= rbytes ? *reinterpret_cast<const v4u*>(e + off) : v4u{};
            }
        }
        for (int64_t p0 = mb; p0 < me; p0 += kMissGroup) {
#pragma unroll
            for (int q = 0; q < kPieces; ++q) *reinterpret_cast<v4u*>(st + lane * kStage + q * 16) = cur[q];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            {
                const uint8_t* e = entry_ptr(p0 + kMissGroup + lane);
                cur[0] = *reinterpret_cast<const v4u*>(e);
#pragma unroll
                for (int q = 1; q < kPieces; ++q) {
                    const int64_t off = 16 + 8 * (int64_t)kb0 + 16 * (q - 1);
                    cur[q] = off + 16 <= rbytes ? *reinterpret_cast<const v4u*>(e + off) : v4u{};
                }
            }
            const int cnt = (me - p0) < kMissGroup ? (int)(me - p0) : kMissGroup;
            for (int jx = 0; jx < cnt; ++jx) {
                const uint8_t* se = st + jx * kStage;
                const uint32_t cw = *reinterpret_cast<const uint32_t*>(se + 4 * cword);
                const double zs = (double)((int32_t)(cw << cshift) >> 30);
                const double* av = reinterpret_cast<const double*>(se + 16);
#pragma unroll
                for (int k = 0; k < KB; ++k) acc[k] = fma(av[k], zs, acc[k]);
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
        if (valid) {
#pragma unroll
            for (int k = 0; k < KB; ++k) {
                const int bin = kb0 + k;
                const int c = bin * v_pad + v;
                double val = acc[k];
                if (first) val += (double)y[(int64_t)c * ldy + smp] * col_scale[c] - cst[c];
                const int mode = (col_mode && first) ? col_mode[bin] : 2;
                double* dst = xz + (int64_t)(bin * v_count + v) * ldxz + smp;
                if (mode == 1) *dst = val;
                else *dst += val;
            }
        }
    }
}

// One bin of k_p1_finalize_z8 without staging: one wave per sample, lane j
// takes the sample's entries j, j + 64, ... of [ra, rb) and keeps all 64
// columns' sums of z * a in registers, so each entry's record (code words and
// the bin's a) reaches one lane instead of being broadcast to 64 through LDS.
// Per (entry, column) one signed bit-field extract, one conversion and one
// FMA.  The wave's 64 x 64 partial sums are then combined by a halving
// exchange (xor 32, 16, ..., 1; at each step a lane keeps the half of its
// values whose column bit equals its lane bit and adds the partner's copy of
// that half), which leaves column c's total in lane c.
__global__ __launch_bounds__(256) void k_p1_finalize_z8_lanes(
    const int64_t* __restrict__ y, int64_t ldy, int64_t n, int n_bins, int kb0, int v_count, int v_pad,
    const double* __restrict__ col_scale, const double* __restrict__ cst, const int64_t* __restrict__ miss_ptr,
    const int32_t* __restrict__ miss_idx, int64_t k0, int64_t ra, int64_t rb, const uint8_t* __restrict__ rec,
    double* __restrict__ xz, int64_t ldxz, const uint8_t* __restrict__ col_mode) {
    const bool first = ra == k0;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t smp = (int64_t)blockIdx.x * 4 + wave;
    const int lane = threadIdx.x & 63;
    if (smp >= n) return;
    const int64_t b0 = miss_ptr[smp], b1 = miss_ptr[smp + 1];
    const int64_t mb = wave_lower_bound(miss_idx, b0, b1, ra, lane);
    const int64_t me = wave_lower_bound(miss_idx, mb, b1, rb, lane);
    const int vw = (v_pad + 63) / 64;
    const int64_t rbytes = miss_rec_bytes(n_bins);
    const int64_t a_off = 16 + 8 * (int64_t)kb0;
    for (int wd = 0; wd < vw; ++wd) {
        const uint8_t* rbase = rec + (int64_t)wd * rbytes;
        double acc[64];
#pragma unroll
        for (int c = 0; c < 64; ++c) acc[c] = 0.0;
        // a lane without an entry holds zero code words and a = 0 (adds nothing)
        v4u h = v4u{};
        double av = 0.0;
        if (mb + lane < me) {
            const uint8_t* e = rbase + ((int64_t)miss_idx[mb + lane] - k0) * vw * rbytes;
            h = *reinterpret_cast<const v4u*>(e);
            av = *reinterpret_cast<const double*>(e + a_off);
        }
        int32_t idx1 = mb + 64 + lane < me ? miss_idx[mb + 64 + lane] : (int32_t)k0;
        // the first entry and the next index have landed before the loop, so the
        // loop header sees no pending load from outside and each iteration waits
        // only where it moves its prefetched entry into place (after the sums)
        __builtin_amdgcn_s_waitcnt(waitcnt_vm(0));
        for (int64_t p0 = mb; p0 < me; p0 += 64) {
            // the lane's next entry (its index loaded one iteration ago) and the
            // index after it are loaded while this entry is summed
            const int64_t pn = p0 + 64 + lane;
            v4u hn = v4u{};
            double an = 0.0;
            if (pn < me) {
                const uint8_t* e = rbase + ((int64_t)idx1 - k0) * vw * rbytes;
                hn = *reinterpret_cast<const v4u*>(e);
                an = *reinterpret_cast<const double*>(e + a_off);
            }
            const int32_t idx2 = pn + 64 < me ? miss_idx[pn + 64] : (int32_t)k0;
#pragma unroll
            for (int c = 0; c < 64; ++c) {
                const int32_t zc = (int32_t)(h[c >> 4] << (30 - 2 * (c & 15))) >> 30;
                acc[c] = fma(av, (double)zc, acc[c]);
            }
            h = hn;
            av = an;
            idx1 = idx2;
        }
#pragma unroll
        for (int sft = 32; sft >= 1; sft >>= 1) {
            const bool up = (lane & sft) != 0;
#pragma unroll
            for (int i = 0; i < sft; ++i) {
                const double keep = up ? acc[i + sft] : acc[i];
                const double send = up ? acc[i] : acc[i + sft];
                acc[i] = keep + __shfl_xor(send, sft, 64);
            }
        }
        const int v = wd * 64 + lane;
        if (v < v_count) {
            const int c = kb0 * v_pad + v;
            double val = acc[0];
            if (first) val += (double)y[(int64_t)c * ldy + smp] * col_scale[c] - cst[c];
            const int mode = (col_mode && first) ? col_mode[kb0] : 2;
            double* dst = xz + (int64_t)(kb0 * v_count + v) * ldxz + smp;
            if (mode == 1) *dst = val;
            else *dst += val;
        }
    }
}

void launch_p1_finalize_z8(const int64_t* y, int64_t ldy, int64_t n, int n_bins, int v_count, int v_pad,
                           const double* col_scale, const double* cst, const int64_t* miss_ptr,
                           const int32_t* miss_idx, int64_t k0, int64_t k1, const void* rec, double* xz,
                           int64_t ldxz, const uint8_t* col_mode, int64_t range_bytes, hipStream_t st) {
    if (n <= 0) return;
    const int vw = (int)ceil_div(v_pad, 64);
    const int64_t rec_stride = (int64_t)vw * miss_rec_bytes(n_bins);
    // SNP ranges of range_bytes of records, one launch per range (and bin
    // group), in order.  Small ranges keep a range L2-resident while all
    // samples walk it.  Measured at N=100,000, M=500,000, B=64: the one-bin
    // kernel (no LDS staging) is fastest at 8 MiB (2.45 ms; 3.1 ms at 16 MiB
    // and up, 2.9 ms at 4 MiB), while the LDS-staged kernel of several bins
    // favoured fewer, larger launches (2 MiB 3.9 ms, 16 MiB and up 3.3 ms)
    if (range_bytes <= 0) range_bytes = n_bins == 1 ? (int64_t)8 << 20 : (int64_t)64 << 20;
    const int64_t range = std::max<int64_t>(4096, range_bytes / rec_stride);
    const uint8_t* r = static_cast<const uint8_t*>(rec);
    for (int64_t ra = k0; ra < k1 || ra == k0; ra += range) {
        const int64_t rb = std::min(k1, ra + range);
        // bins in groups of 16, 8, 4, 2 (staged) and 1 (k_p1_finalize_z8_lanes)
        for (int kb0 = 0; kb0 < n_bins;) {
            const int left = n_bins - kb0;
            const dim3 grid((unsigned)ceil_div(n, 4));
#define GPE_FZ8(KB_)                                                                                             \
    hipLaunchKernelGGL(k_p1_finalize_z8<KB_>, grid, dim3(256), 0, st, y, ldy, n, n_bins, kb0, v_count, v_pad,    \
                       col_scale, cst, miss_ptr, miss_idx, k0, ra, rb, r, xz, ldxz, col_mode);                    \
    kb0 += KB_
            if (left >= 16) {
                GPE_FZ8(16);
            } else if (left >= 8) {
                GPE_FZ8(8);
            } else if (left >= 4) {
                GPE_FZ8(4);
            } else if (left >= 2) {
                GPE_FZ8(2);
            } else {
                hipLaunchKernelGGL(k_p1_finalize_z8_lanes, grid, dim3(256), 0, st, y, ldy, n, n_bins, kb0, v_count,
                                   v_pad, col_scale, cst, miss_ptr, miss_idx, k0, ra, rb, r, xz, ldxz, col_mode);
                kb0 += 1;
            }
#undef GPE_FZ8
            GPE_HIP(hipGetLastError());
        }
        if (rb >= k1) break;
    }
}

// column max |x| of a column-major FP64 matrix whose column bin*V+v maps to padded column bin*v_pad+v
__global__ void k_colmax_f64(const double* x, int64_t ldx, int64_t n, int v_count, int v_pad,
                             uint64_t* colmax_bits) {
    const int c = blockIdx.y;
    const int k = c / v_pad, v = c % v_pad;
    if (v >= v_count) return;
    const double* col = x + (int64_t)(k * v_count + v) * ldx;
    double mx = 0.0;
    int bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = fabs(col[i]);
        if (a <= DBL_MAX)
            mx = fmax(mx, a);
        else
            bad = 1;  // NaN or infinity
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 64));
    bad = __any(bad);
    // a positive quiet NaN's bits exceed every finite value's: the column stays marked
    if ((threadIdx.x & 63) == 0 && (bad || mx > 0.0))
        atomicMax(reinterpret_cast<unsigned long long*>(colmax_bits + c),
                  bad ? 0x7FF8000000000000ULL : (unsigned long long)__double_as_longlong(mx));
}
void launch_colmax_f64(const double* x, int64_t ldx, int64_t n, int n_cols, int v_count, int v_pad,
                       uint64_t* colmax_bits, hipStream_t st) {
    if (n <= 0 || n_cols <= 0) return;
    int64_t gx = ceil_div(n, 256);
    if (gx > 64) gx = 64;
    hipLaunchKernelGGL(k_colmax_f64, dim3((unsigned)gx, (unsigned)n_cols), dim3(256), 0, st, x, ldx, n, v_count,
                       v_pad, colmax_bits);
    GPE_HIP(hipGetLastError());
}

// Phase-2 operand: limbs of round(Xz * 2^shift) over samples, plus exact column sums.
// Swizzled limbs of the phase-2 operand x (row c' = bin * v_count + v of length n,
// leading dimension ldx).  One workgroup per 64-column x 64-sample tile: the
// quantized values q = round(x 2^shift_c) (0 past n or v_count) are read with
// threads over samples (coalesced) into LDS; then thread (column, group of 16
// samples) adds its 16 values to the column's exact sum and stores each limb's
// 16 digits as one 16-byte operand unit in the transposed byte order of the
// unpacked genotype operand (threads over columns: consecutive units).
template <int NL>
__global__ __launch_bounds__(256) void k_quant_p2_t(const double* x, int64_t ldx, int64_t n, int v_count, int v_pad,
                                                    int ncols, const int32_t* shift, int ctc, int64_t coef_ctiles,
                                                    int8_t* coef, int64_t* col_sum) {
    __shared__ int64_t qt[64][65];
    const int tid = threadIdx.x;
    const int64_t n0 = (int64_t)blockIdx.x * 64;
    const int c0 = (int)blockIdx.y * 64;
    for (int idx = tid; idx < 64 * 64; idx += 256) {
        const int cl = idx >> 6, sl = idx & 63;
        const int c = c0 + cl;
        const int64_t smp = n0 + sl;
        int64_t q = 0;
        if (c < ncols && smp < n) {
            const int k = c / v_pad, v = c % v_pad;
            if (v < v_count) q = quant_rn(x[(int64_t)(k * v_count + v) * ldx + smp], shift[c]);
        }
        qt[cl][sl] = q;
    }
    __syncthreads();
    const int cl = tid & 63, gq = tid >> 6;
    const int c = c0 + cl;
    if (c >= ncols) return;
    uint32_t unit[NL][4];
#pragma unroll
    for (int l = 0; l < NL; ++l) unit[l][0] = unit[l][1] = unit[l][2] = unit[l][3] = 0u;
    int64_t qs = 0;
#pragma unroll 4
    for (int j = 0; j < 16; ++j) {
        const int64_t q = qt[cl][gq * 16 + j];
        qs += q;
        int8_t d[NL];
        balanced_digits<NL>(q, d);
        const int b = ((j & 3) << 2) | (j >> 2);  // the unpacked operand's byte order (swz_offset)
#pragma unroll
        for (int l = 0; l < NL; ++l) unit[l][b >> 2] |= (uint32_t)(uint8_t)d[l] << (8 * (b & 3));
    }
    if (qs != 0) atomicAdd(reinterpret_cast<unsigned long long*>(col_sum + c), (unsigned long long)qs);
    const int64_t s16 = n0 + gq * 16;  // first sample of the unit
    const int64_t kl = s16 / kKChunk;
    const int kk = (int)(s16 % kKChunk);
    const int ct = c >> 5, lr = c & 31;
    const int cgp = ct / ctc, ctcl = ct % ctc;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const int64_t tg = (int64_t)cgp * (NL * ctc) + l * ctc + ctcl;
        *reinterpret_cast<v4u*>(coef + swz_offset(kl, coef_ctiles, tg, kk, lr)) =
            v4u{unit[l][0], unit[l][1], unit[l][2], unit[l][3]};
    }
}
void launch_quant_p2(const double* x, int64_t ldx, int64_t n, int n_bins, int v_count, int v_pad,
                     const int32_t* shift, int nl, int ctc, int64_t coef_ctiles, int8_t* coef, int64_t* col_sum,
                     hipStream_t st) {
    const int64_t kcount = ceil_div(n, kKChunk);
    const int ncols = n_bins * v_pad;
    const dim3 grid((unsigned)(kcount * (kKChunk / 64)), (unsigned)ceil_div(ncols, 64));
    if (nl == 5)
        hipLaunchKernelGGL(k_quant_p2_t<5>, grid, dim3(256), 0, st, x, ldx, n, v_count, v_pad, ncols, shift, ctc,
                           coef_ctiles, coef, col_sum);
    else if (nl == 6)
        hipLaunchKernelGGL(k_quant_p2_t<6>, grid, dim3(256), 0, st, x, ldx, n, v_count, v_pad, ncols, shift, ctc,
                           coef_ctiles, coef, col_sum);
    else
        throw Error("phase-2 operand limbs must be 5 or 6");
    GPE_HIP(hipGetLastError());
}

// Column sums supplied by the caller in place of the exact sums of the quantized
// operand: col_sum[c] = round(sum[k * v_count + v] 2^shift_c) for the padded
// column c = k * v_pad + v (the centring term of the reference's Mailman phase 2,
// src/native/gwldcore.cpp:2965-2967, :874-880).  A non-finite sum makes the
// column's scale NaN, as the reference's product with it is NaN.
__global__ void k_col_sum_override(const double* sum, int v_count, int v_pad, int ncols, const int32_t* shift,
                                   double* col_scale, int64_t* col_sum) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncols) return;
    const int k = c / v_pad, v = c % v_pad;
    if (v >= v_count) return;
    const double s = sum[k * v_count + v];
    const double t = ldexp(s, shift[c]);
    if (fabs(t) < 4.0e18) {
        col_sum[c] = __double2ll_rn(t);
    } else {
        col_sum[c] = 0;
        col_scale[c] = __builtin_nan("");
    }
}
void launch_col_sum_override(const double* sum, int n_bins, int v_count, int v_pad, const int32_t* shift,
                             double* col_scale, int64_t* col_sum, hipStream_t st) {
    const int ncols = n_bins * v_pad;
    if (ncols <= 0) return;
    hipLaunchKernelGGL(k_col_sum_override, dim3((unsigned)ceil_div(ncols, 256)), dim3(256), 0, st, sum, v_count,
                       v_pad, ncols, shift, col_scale, col_sum);
    GPE_HIP(hipGetLastError());
}

// Flag int8 probe values outside {-1, 0, +1} in rows x cols of z (leading dimension ldz).
__global__ void k_z8_check(const int8_t* z, int64_t ldz, int64_t rows, int cols, unsigned int* bad) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    const int8_t v = z[(idx / cols) * ldz + idx % cols];
    if (v < -1 || v > 1) atomicOr(bad, 1u);
}
void launch_z8_check(const int8_t* z, int64_t ldz, int64_t rows, int cols, unsigned int* bad, hipStream_t st) {
    if (rows <= 0 || cols <= 0) return;
    hipLaunchKernelGGL(k_z8_check, dim3((unsigned)ceil_div(rows * cols, 256)), dim3(256), 0, st, z, ldz, rows, cols,
                       bad);
    GPE_HIP(hipGetLastError());
}

// Limb rows of the phase-2 operand for gathers by sample.  One workgroup per
// 64-sample x 64-column tile: q and its balanced digits as k_quant_p2_t forms
// them, staged in LDS and written out so that each sample's nl x 64 bytes are
// contiguous (byte 64 l + c: digit l, most significant first, of column c).
// Rows run to n inclusive: row n is all zeros (the pad entry of the gathers).
template <int NL>
__global__ __launch_bounds__(256) void k_xrow_limbs(const double* __restrict__ x, int64_t ldx, int64_t n, int ncols,
                                                    int v_count, int v_pad, const int32_t* __restrict__ shift,
                                                    int nch, int8_t* __restrict__ xrow) {
    constexpr int DW = NL * 16;  // dwords per (sample, chunk)
    __shared__ __attribute__((aligned(16))) int8_t tile[64][NL * 64];
    const int tid = threadIdx.x;
    const int64_t n0 = (int64_t)blockIdx.x * 64;
    const int h = blockIdx.y;
    for (int i = tid; i < 64 * 64; i += 256) {
        const int cc = i >> 6, sl = i & 63;
        const int c = h * 64 + cc;
        const int64_t smp = n0 + sl;
        int64_t q = 0;
        if (c < ncols && smp < n) {
            const int k = c / v_pad, v = c % v_pad;
            if (v < v_count) q = quant_rn(x[(int64_t)(k * v_count + v) * ldx + smp], shift[c]);
        }
        int8_t d[NL];
        balanced_digits<NL>(q, d);
#pragma unroll
        for (int l = 0; l < NL; ++l) tile[sl][l * 64 + cc] = d[l];
    }
    __syncthreads();
    for (int i = tid; i < 64 * DW; i += 256) {
        const int sl = i / DW, dw = i % DW;
        const int64_t smp = n0 + sl;
        if (smp <= n)
            reinterpret_cast<uint32_t*>(xrow)[(smp * nch + h) * DW + dw] =
                reinterpret_cast<const uint32_t*>(tile[sl])[dw];
    }
}
void launch_xrow_limbs(const double* x, int64_t ldx, int64_t n, int n_bins, int v_count, int v_pad,
                       const int32_t* shift, int nl, int nch, int8_t* xrow, hipStream_t st) {
    if (n <= 0) return;
    const dim3 grid((unsigned)ceil_div(n + 1, 64), (unsigned)nch);
    if (nl == 5)
        hipLaunchKernelGGL(k_xrow_limbs<5>, grid, dim3(256), 0, st, x, ldx, n, n_bins * v_pad, v_count, v_pad, shift,
                           nch, xrow);
    else if (nl == 6)
        hipLaunchKernelGGL(k_xrow_limbs<6>, grid, dim3(256), 0, st, x, ldx, n, n_bins * v_pad, v_count, v_pad, shift,
                           nch, xrow);
    else
        throw Error("phase-2 operand limbs must be 5 or 6");
    GPE_HIP(hipGetLastError());
}

// rptr[j * (P + 1) + p] = first entry of SNP j's (sorted) missing list whose
// sample is >= p R, for p = 0..P: the list's part in sample range p.
__global__ void k_range_ptrs(const int64_t* __restrict__ ptr, const int32_t* __restrict__ idx, int64_t m, int parts,
                             int64_t range, int64_t* __restrict__ rptr) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= m * (parts + 1)) return;
    const int64_t j = t / (parts + 1);
    const int64_t bound = (t % (parts + 1)) * range;
    int64_t lo = ptr[j], hi = ptr[j + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (idx[mid] < bound) lo = mid + 1;
        else hi = mid;
    }
    rptr[t] = lo;
}
void launch_range_ptrs(const int64_t* ptr, const int32_t* idx, int64_t m, int parts, int64_t range, int64_t* rptr,
                       hipStream_t st) {
    const int64_t total = m * (parts + 1);
    if (total <= 0) return;
    hipLaunchKernelGGL(k_range_ptrs, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, ptr, idx, m, parts, range,
                       rptr);
    GPE_HIP(hipGetLastError());
}

// Exact missing-call sums of the phase-2 operand:
//   out[j - r0][64 h + cc] += sum over samples n missing at j of q[n][64 h + cc]
// (out zeroed first).  One wave per (sample range p, column chunk h, SNP j); the
// ranges' int64 sums meet in out by atomic adds, exact in any order.
// A sample's limb row (k_xrow_limbs) is P = 4 NL pieces of 16 bytes.  The sums
// run on the int8 matrix cores: v_mfma_i32_16x16x64_i8 with A the constant
// selector A[i][k] = 1 where k is byte i of a lane's 16 (byte (lane % 16) of
// each lane set) gives D[i][c] += sum over the four lanes c, c+16, c+32, c+48
// of byte i of their B registers -- the same (row, column) entry of D for the
// same byte of four different gathered pieces, whatever order the instruction
// gives the K index within a lane (A and B share it).  So, per round of 16
// entries, lane class c = lane % 16 loads piece c of entries 4m + lane / 16 in
// four MFMAs into D0 (D0[i][c]: byte 16 c + i of the row), and the T = P - 16
// tail pieces of all 16 entries go through T / 4 more MFMAs into D1, class c
// loading piece 16 + c % T (D1[i][c]: byte 16 (16 + c % T) + i, summed over the
// 16 / T classes of that piece at the end).  Per entry that is 5/16 (NL = 5) of
// an MFMA and the row's 20 lanes of 16-byte loads, where summing bytes on the
// vector ALU took about 8 instructions per entry; the limbs are combined to
// int64 at the end (|sum| <= 128 n < 2^31 per digit).  A batch's entries come 64
// at a time (lane t loads entry t's sample index; entries past the list read the
// all-zero row n).  With P a multiple of the XCD count, workgroup b runs on XCD
// b % 8 and the ranges are dealt so that each XCD gathers only from its own
// ranges, one after another: a range's rows stay in that XCD's L2.
// BUF (table below 4 GiB): buffer loads with 32-bit offsets.
template <int NL, bool BUF>
__global__ __launch_bounds__(256) void k_miss_sums(const int64_t* __restrict__ rptr, int parts,
                                                   const int32_t* __restrict__ idx, int64_t r0, int64_t r1,
                                                   const int8_t* __restrict__ xrow, int n, int nch,
                                                   int64_t* __restrict__ out, int64_t ld_ms, uint32_t table_bytes) {
    constexpr int P = 4 * NL;      // 16-byte pieces per row
    constexpr int T = P - 16;      // tail pieces
    constexpr int R = 16 / T;      // lane classes per tail piece
    constexpr int TM = T / 4;      // tail MFMAs per round of 16 entries
    constexpr int TE = 64 / T;     // entries per tail MFMA
    // rounds of 16 entries per step: occupancy beats loads in flight per wave
    // (N = 100,000, M = 500,000, 64 columns, 0.5% missing: 1 round 4.78 ms with
    // the range sums, 2 rounds 4.86 ms, 4 rounds 6.34 ms)
    constexpr int RPS = 1;
    static_assert(T == 4 || T == 8, "limb rows of 5 or 6 limbs");
    __shared__ int32_t red[4][512];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int64_t rows = r1 - r0;
    const int64_t per_range = rows * nch;
    int p;
    int64_t item;
    if (parts % kXcds == 0) {
        const int64_t br = (per_range + 3) / 4;  // workgroups per range
        const int64_t i = blockIdx.x / kXcds;
        const int q = (int)(i / br);
        p = (int)(blockIdx.x % kXcds) + kXcds * q;
        item = (i % br) * 4 + wave;
        if (q >= parts / kXcds || item >= per_range) return;
    } else {
        const int64_t w = (int64_t)blockIdx.x * 4 + wave;
        p = (int)(w / per_range);
        item = w % per_range;
        if (p >= parts) return;
    }
    const int h = (int)(item / rows);
    const int64_t j = r0 + item % rows;
    const int64_t b0 = rptr[j * (parts + 1) + p], b1 = rptr[j * (parts + 1) + p + 1];
    const uint32_t rstride = (uint32_t)nch * (uint32_t)(NL * 64);  // bytes per sample row
    const int8_t* base = xrow + (int64_t)h * (NL * 64);
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<int8_t*>(base), (short)0, (int)(table_bytes - (uint32_t)h * (uint32_t)(NL * 64)), 0x00020000);
    const int c16 = lane & 15, g = lane >> 4;
    // the selector: byte c16 of the lane's 16 is 1
    const int32_t one = 1 << (8 * (c16 & 3));
    const v4i sel = v4i{(c16 >> 2) == 0 ? one : 0, (c16 >> 2) == 1 ? one : 0, (c16 >> 2) == 2 ? one : 0,
                        (c16 >> 2) == 3 ? one : 0};
    const uint32_t off_head = 16u * (uint32_t)c16;
    const uint32_t off_tail = 16u * (uint32_t)(16 + c16 % T);
    const int slot_tail = g * R + c16 / T;
    auto load = [&](uint32_t smp, uint32_t off) -> v4i {
        if constexpr (BUF)
            return __builtin_bit_cast(v4i, __builtin_amdgcn_raw_buffer_load_b128(rsrc, (int)(smp * rstride + off), 0, 0));
        else
            return *reinterpret_cast<const v4i*>(base + (int64_t)smp * rstride + off);
    };
    v4i d0a = v4i{0, 0, 0, 0}, d0b = v4i{0, 0, 0, 0}, d1 = v4i{0, 0, 0, 0};
    for (int64_t p0 = b0; p0 < b1; p0 += 64) {
        const int cnt = (int)((b1 - p0) < 64 ? (b1 - p0) : 64);
        const int my = lane < cnt ? idx[p0 + lane] : n;
        // RPS rounds of 16 entries per step: RPS (4 + TM) loads in flight
        for (int t0 = 0; t0 < cnt; t0 += 16 * RPS) {
            v4i hv[4 * RPS], tv[RPS * TM];
#pragma unroll
            for (int r = 0; r < RPS; ++r) {
#pragma unroll
                for (int m = 0; m < 4; ++m)
                    hv[4 * r + m] = load((uint32_t)__shfl(my, t0 + 16 * r + 4 * m + g, 64), off_head);
#pragma unroll
                for (int m = 0; m < TM; ++m)
                    tv[TM * r + m] = load((uint32_t)__shfl(my, t0 + 16 * r + TE * m + slot_tail, 64), off_tail);
            }
            // every load issued before the first MFMA waits on one
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int m = 0; m < 4 * RPS; m += 2) {
                d0a = __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, hv[m], d0a, 0, 0, 0);
                d0b = __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, hv[m + 1], d0b, 0, 0, 0);
            }
#pragma unroll
            for (int m = 0; m < RPS * TM; ++m) d1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(sel, tv[m], d1, 0, 0, 0);
        }
    }
    // D[i][c] sits in lane c + 16 (i / 4), register i % 4
    int32_t* rd = red[wave];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        rd[(4 * g + r) * 16 + c16] = d0a[r] + d0b[r];
        rd[256 + (4 * g + r) * 16 + c16] = d1[r];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // lane = column: digit l is row byte 64 l + lane
    int64_t me = 0;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const int bb = 64 * l + lane;
        int32_t v = 0;
        if (l < 4) {
            v = rd[(bb & 15) * 16 + (bb >> 4)];
        } else {
            const int q = (bb >> 4) - 16;
#pragma unroll
            for (int a = 0; a < R; ++a) v += rd[256 + (bb & 15) * 16 + a * T + q];
        }
        me = me * 256 + (int64_t)v;
    }
    // every sample range adds into the one zeroed sum: integer atomics, exact in any order
    atomicAdd(reinterpret_cast<unsigned long long*>(&out[(j - r0) * ld_ms + h * 64 + lane]),
              (unsigned long long)me);
}
// Per-sample sums of squared standardized genotypes weighted per SNP and bin,
// the X2 * W panel products of the fused residual-variance pass
// (src/native/gwldcore.cpp:3496-3521):
//   part[y][k][i] = sum over SNPs j of chunk y of xt_ij^2 w[(j - s) + (kb0 + k) ldw]
// with xt_ij = x_ij - sum_a C[i + a ldc] rx[(j - s) ldrx + a] (no projection when
// p = 0) and x_ij the standardized genotype of the working codes (0 at a missing
// call under mean imputation).  One thread per sample: its row of the
// sample-major layout is read 16 bytes (64 SNPs) at a time, and the SNP-wise
// values (mean, 1/sd, weights, R x) are the same for every lane of a wave.
template <int KB, int PMAX>
__global__ __launch_bounds__(256) void k_sq_rowsums(const uint8_t* __restrict__ smp, int64_t stride, int64_t n,
                                                    int64_t s, int64_t e, int64_t chunk,
                                                    const double* __restrict__ mean,
                                                    const double* __restrict__ inv_sd, const double* __restrict__ w,
                                                    int64_t ldw, int kb0, int kb, const double* __restrict__ C,
                                                    int64_t ldc, const double* __restrict__ rx, int64_t ldrx, int p,
                                                    double* __restrict__ part, int64_t ldp, int64_t part_stride) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t j0 = s + (int64_t)blockIdx.y * chunk;
    const int64_t j1 = j0 + chunk < e ? j0 + chunk : e;
    if (i >= n) return;
    double ci[PMAX > 0 ? PMAX : 1];
#pragma unroll
    for (int a = 0; a < PMAX; ++a) ci[a] = a < p ? C[(int64_t)a * ldc + i] : 0.0;
    double acc[KB];
#pragma unroll
    for (int k = 0; k < KB; ++k) acc[k] = 0.0;
    const uint8_t* row = smp + i * stride;
    for (int64_t jb = j0 & ~(int64_t)63; jb < j1; jb += 64) {
        const v4u wv = *reinterpret_cast<const v4u*>(row + jb / 4);
        const int t0 = jb < j0 ? (int)(j0 - jb) : 0;
        const int t1 = jb + 64 > j1 ? (int)(j1 - jb) : 64;
        for (int t = t0; t < t1; ++t) {
            const int64_t j = jb + t;
            const uint32_t code = (wv[t >> 4] >> (2 * (t & 15))) & 3u;
            double x = code == 1u ? 0.0 : ((double)(code == 0u ? 0 : (int)code - 1) - mean[j]) * inv_sd[j];
#pragma unroll
            for (int a = 0; a < PMAX; ++a)
                if (a < p) x -= ci[a] * rx[(j - s) * ldrx + a];
            const double x2 = x * x;
#pragma unroll
            for (int k = 0; k < KB; ++k)
                if (k < kb) acc[k] = fma(x2, w[(j - s) + (int64_t)(kb0 + k) * ldw], acc[k]);
        }
    }
#pragma unroll
    for (int k = 0; k < KB; ++k)
        if (k < kb) part[(int64_t)blockIdx.y * part_stride + (int64_t)k * ldp + i] = acc[k];
}
// out[(kb0 + k) ldo + i] += sum_y part[y][k][i], the chunks added in order
__global__ void k_sq_rowsums_reduce(const double* __restrict__ part, int ny, int64_t part_stride, int64_t ldp,
                                    int64_t n, int kb0, int kb, double* __restrict__ out, int64_t ldo) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * kb) return;
    const int k = (int)(t / n);
    const int64_t i = t % n;
    double v = 0.0;
    for (int y = 0; y < ny; ++y) v += part[(int64_t)y * part_stride + (int64_t)k * ldp + i];
    out[(int64_t)(kb0 + k) * ldo + i] += v;
}
int64_t sq_rowsums_chunk(int64_t n, int64_t L) {
    // about 2048 workgroups, chunks of at least 1024 SNPs
    const int64_t gx = ceil_div(n, 256);
    const int64_t ny = std::max<int64_t>(1, std::min<int64_t>(ceil_div(L, 1024), ceil_div(2048, gx)));
    return round_up(ceil_div(L, ny), 64);
}
void launch_sq_rowsums(const uint8_t* smp, int64_t stride, int64_t n, int64_t s, int64_t e, const double* mean,
                       const double* inv_sd, const double* w, int64_t ldw, int n_bins, const double* C, int64_t ldc,
                       const double* rx, int64_t ldrx, int p, double* part, double* out, int64_t ldo, hipStream_t st) {
    const int64_t L = e - s;
    if (n <= 0 || L <= 0 || n_bins <= 0) return;
    const int64_t chunk = sq_rowsums_chunk(n, L);
    const int64_t ny = ceil_div(L, chunk);
    const int64_t ldp = n;
    const int64_t part_stride = 8 * n;
    const dim3 grid((unsigned)ceil_div(n, 256), (unsigned)ny);
    for (int kb0 = 0; kb0 < n_bins; kb0 += 8) {
        const int kb = std::min(8, n_bins - kb0);
#define GPE_SQ(KB_, P_)                                                                                          \
    hipLaunchKernelGGL((k_sq_rowsums<KB_, P_>), grid, dim3(256), 0, st, smp, stride, n, s, e, chunk, mean, inv_sd, \
                       w, ldw, kb0, kb, C, ldc, rx, ldrx, p, part, ldp, part_stride)
#define GPE_SQP(KB_)                     \
    if (p == 0) GPE_SQ(KB_, 0);          \
    else if (p <= 8) GPE_SQ(KB_, 8);     \
    else if (p <= 32) GPE_SQ(KB_, 32);   \
    else GPE_SQ(KB_, 64)
        if (kb == 1) {
            GPE_SQP(1);
        } else if (kb == 2) {
            GPE_SQP(2);
        } else if (kb <= 4) {
            GPE_SQP(4);
        } else {
            GPE_SQP(8);
        }
#undef GPE_SQP
#undef GPE_SQ
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sq_rowsums_reduce, dim3((unsigned)ceil_div(n * kb, 256)), dim3(256), 0, st, part,
                           (int)ny, part_stride, ldp, n, kb0, kb, out, ldo);
        GPE_HIP(hipGetLastError());
    }
}

void launch_miss_sums(const int64_t* rptr, int parts, const int32_t* idx, int64_t r0, int64_t r1, const int8_t* xrow,
                      int64_t n, int nl, int nch, int64_t* out, int64_t ld_ms, hipStream_t st) {
    const int64_t per_range = (r1 - r0) * nch;
    if (per_range <= 0) return;
    if ((n + 1) * nch * nl * 16 >= ((int64_t)1 << 32)) throw Error("missing-call gather table exceeds 2^32 dwords");
    const int64_t blocks = parts % kXcds == 0 ? kXcds * (int64_t)(parts / kXcds) * ceil_div(per_range, 4)
                                              : ceil_div(per_range * parts, 4);
    const dim3 grid((unsigned)blocks);
    const int64_t table_bytes = (n + 1) * nch * nl * 64;
    const bool buf = table_bytes <= (int64_t)0xFFFFFFFF;
    const uint32_t tb = buf ? (uint32_t)table_bytes : 0u;
#define GPE_MS(NL_, BUF_)                                                                                           \
    hipLaunchKernelGGL((k_miss_sums<NL_, BUF_>), grid, dim3(256), 0, st, rptr, parts, idx, r0, r1, xrow, (int)n,  \
                       nch, out, ld_ms, tb)
    if (nl == 5 && buf)
        GPE_MS(5, true);
    else if (nl == 5)
        GPE_MS(5, false);
    else if (nl == 6 && buf)
        GPE_MS(6, true);
    else if (nl == 6)
        GPE_MS(6, false);
    else
        throw Error("phase-2 operand limbs must be 5 or 6");
#undef GPE_MS
    GPE_HIP(hipGetLastError());
}

// Phase-2 finalize (gwldcore.cpp:2525-2593): per (SNP, bin) sum the per-tile
// Work^2 sums, scale by inv^2/(N_denom-1)^2, guard, and accumulate.
__global__ void k_p2_finalize(const double* s2, const double* s4, int64_t ld_s, int64_t r0, int64_t r1, int n_bins,
                              int tiles_per_bin, int v_count, const double* inv_left, double inv_denom2,
                              double extreme_warn, int drop_extreme, double* meansq, int64_t ld_meansq,
                              double* mc_agg, double* mc_per_snp, unsigned long long* guard_counts) {
    const int64_t j = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= r1) return;
    double acc = 0.0, acc4 = 0.0;
    for (int t = 0; t < tiles_per_bin; ++t) {
        acc += s2[j * ld_s + k * tiles_per_bin + t];
        if (s4) acc4 += s4[j * ld_s + k * tiles_per_bin + t];
    }
    const double inv = inv_left[j - r0];
    const double left = (isfinite(inv) && inv > 0.0) ? inv * inv * inv_denom2 : 0.0;
    double upd = acc * left;
    const double seg = (double)v_count;
    if (!isfinite(upd)) {
        atomicAdd(&guard_counts[0], 1ULL);
        return;
    }
    if (isfinite(extreme_warn) && seg > 0.0 && upd > extreme_warn * seg) {
        atomicAdd(&guard_counts[1], 1ULL);
        if (drop_extreme) return;
    }
    if (upd < 0.0) upd = 0.0;
    meansq[j * ld_meansq + k] += upd;
    if (s4) {
        const double sumsq = acc4 * left * left;
        if (mc_per_snp) mc_per_snp[j * ld_meansq + k] += sumsq;
        if (mc_agg) atomicAdd(&mc_agg[k], sumsq);
    }
}
void launch_p2_finalize(const double* s2, const double* s4, int64_t ld_s, int64_t r0, int64_t r1, int n_bins,
                        int tiles_per_bin, int v_count, const double* inv_left, double inv_denom2,
                        double extreme_warn, int drop_extreme, double* meansq, int64_t ld_meansq, double* mc_agg,
                        double* mc_per_snp, unsigned long long* guard_counts, hipStream_t st) {
    if (r1 <= r0) return;
    hipLaunchKernelGGL(k_p2_finalize, dim3((unsigned)ceil_div(r1 - r0, 256), (unsigned)n_bins), dim3(256), 0, st, s2,
                       s4, ld_s, r0, r1, n_bins, tiles_per_bin, v_count, inv_left, inv_denom2, extreme_warn,
                       drop_extreme, meansq, ld_meansq, mc_agg, mc_per_snp, guard_counts);
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Exact integer path operands
// ---------------------------------------------------------------------------
// int8 Z (row i, column b; row-major, ldz) -> one limb, two column tiles per group
__global__ void k_coef_i8(const int8_t* z, int64_t ldz, int64_t k0, int64_t k1, int n_cols, int64_t coef_ctiles,
                          int64_t kc0, int64_t kcount, int8_t* coef) {
    const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (kk >= kcount * kKChunk) return;
    const int64_t i = kc0 * kKChunk + kk;
    int8_t val = 0;
    if (c < n_cols && i >= k0 && i < k1) val = z[(i - k0) * ldz + c];
    const int ct = c >> 5, lr = c & 31;
    coef[swz_offset(kk / kKChunk, coef_ctiles, ct, (int)(kk % kKChunk), lr)] = val;
}
void launch_coef_i8(const int8_t* z, int64_t ldz, int64_t k0, int64_t k1, int n_cols, int64_t coef_ctiles,
                    int8_t* coef, hipStream_t st) {
    const int64_t kc0 = k0 / kKChunk;
    const int64_t kcount = ceil_div(k1, kKChunk) - kc0;
    if (kcount <= 0) return;
    dim3 grid((unsigned)ceil_div(kcount * kKChunk, 256), (unsigned)(coef_ctiles * 32));
    hipLaunchKernelGGL(k_coef_i8, grid, dim3(256), 0, st, z, ldz, k0, k1, n_cols, coef_ctiles, kc0, kcount, coef);
    GPE_HIP(hipGetLastError());
}

// int32 U (column b, row n; column-major, ldu) -> NL balanced limbs over samples
template <int NL>
__global__ void k_quant_i32(const int32_t* u, int64_t ldu, int64_t n, int n_cols, int ctc, int64_t coef_ctiles,
                            int64_t kcount, int8_t* coef) {
    const int64_t kk = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int c = blockIdx.y;
    if (kk >= kcount * kKChunk) return;
    int64_t q = 0;
    if (c < n_cols && kk < n) q = u[(int64_t)c * ldu + kk];
    int8_t d[NL];
    balanced_digits<NL>(q, d);
    const int ct = c >> 5, lr = c & 31;
    const int cgp = ct / ctc, ctcl = ct % ctc;
#pragma unroll
    for (int l = 0; l < NL; ++l) {
        const int64_t tg = (int64_t)cgp * (NL * ctc) + l * ctc + ctcl;
        coef[swz_offset(kk / kKChunk, coef_ctiles, tg, (int)(kk % kKChunk), lr)] = d[l];
    }
}
void launch_quant_i32(const int32_t* u, int64_t ldu, int64_t n, int n_cols, int nl, int ctc, int64_t coef_ctiles,
                      int8_t* coef, hipStream_t st) {
    const int64_t kcount = ceil_div(n, kKChunk);
    const int ncols_pad = (int)(coef_ctiles / nl) * 32;
    dim3 grid((unsigned)ceil_div(kcount * kKChunk, 256), (unsigned)ncols_pad);
    if (nl == 3)
        hipLaunchKernelGGL(k_quant_i32<3>, grid, dim3(256), 0, st, u, ldu, n, n_cols, ctc, coef_ctiles, kcount, coef);
    else
        throw Error("int32 operand limbs must be 3");
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Banded exact Gram for windowed LD scores (winldcore.cpp:480-594).
//
// One workgroup per 64x64 SNP tile pair of the band; wave (wi, wj) computes a
// 32x32 block of pairs.  Both MFMA operands are packed SNP rows expanded to
// int8 genotype counts g and missing flags q, so the four exact integer
// products GG, Gq, qG, qq give the centred cross product
//   sum_n (g_i - mu_i o_i)(g_j - mu_j o_j),  o = 1 - q,
// from which r_ij = s_i s_j dot / N with s = sqrt(N / M2) (the ddof-0
// standardization followed by the restandardization over all N rows), the
// unbiased r^2 - (1 - r^2)/(N - 2), and the window mask.  Row / column sums
// against the annotations are reduced through LDS and added atomically.
// ---------------------------------------------------------------------------


template <bool MISS, bool COV>
__global__ __launch_bounds__(256) void k_gram_band(GramArgs a) {
    __shared__ double r2t[4][32][33];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, lr = lane & 31;
    const int wi = wave >> 1, wj = wave & 1;
    const int ti = a.pairs[3 * blockIdx.x], tj = a.pairs[3 * blockIdx.x + 1], ch = a.pairs[3 * blockIdx.x + 2];
    const int64_t cb = a.chr[2 * ch], ce = a.chr[2 * ch + 1];
    const int64_t i0 = cb + (int64_t)ti * 64 + wi * 32;
    const int64_t j0 = cb + (int64_t)tj * 64 + wj * 32;
    const uint8_t* pa = a.packed + (i0 + lr) * a.stride + half * 16;
    const uint8_t* pb = a.packed + (j0 + lr) * a.stride + half * 16;
    v16i gg = v16i{}, gq = v16i{}, qg = v16i{}, qq = v16i{};
    v4i na = *reinterpret_cast<const v4i*>(pa), nb = *reinterpret_cast<const v4i*>(pb);
    for (int64_t kc = 0; kc < a.kc1; ++kc) {
        const v4i wa = na, wb = nb;
        if (kc + 1 < a.kc1) {
            na = *reinterpret_cast<const v4i*>(pa + (kc + 1) * 32);
            nb = *reinterpret_cast<const v4i*>(pb + (kc + 1) * 32);
        }
#pragma unroll
        for (int s = 0; s < 4; ++s) {
            const uint32_t xa = (uint32_t)wa[s];
            const uint32_t xb = (uint32_t)wb[s];
            v4i ga, qa, gb, qb;
            unpack16(xa, ga, qa);
            unpack16(xb, gb, qb);
            gg = __builtin_amdgcn_mfma_i32_32x32x32_i8(ga, gb, gg, 0, 0, 0);
            if (MISS) {
                gq = __builtin_amdgcn_mfma_i32_32x32x32_i8(ga, qb, gq, 0, 0, 0);
                qg = __builtin_amdgcn_mfma_i32_32x32x32_i8(qa, gb, qg, 0, 0, 0);
                qq = __builtin_amdgcn_mfma_i32_32x32x32_i8(qa, qb, qq, 0, 0, 0);
            }
        }
    }
    const double dn = (double)a.n;
    const int64_t j = j0 + lr;
    const int64_t jr = j - a.row0;
    const bool jv = j < ce;
    const double mu_j = jv ? a.mean[j] : 0.0, s_j = jv ? a.scale[j] : 0.0;
    const double sum_j = jv ? (double)a.sum[j] : 0.0;
    const double nm_j = jv ? dn - (double)a.nobs[j] : 0.0;
#pragma unroll
    for (int e = 0; e < 16; ++e) {
        const int il = 8 * (e >> 2) + 4 * half + (e & 3);
        const int64_t i = i0 + il;
        const int64_t ir = i - a.row0;
        double v = 0.0;
        if (jv && i < ce) {
            bool inwin;
            if (a.window_snps > 0) {
                const int64_t d = j > i ? j - i : i - j;
                inwin = d <= a.window_snps;
            } else {
                const double d = fabs((double)a.bp[ir] - (double)a.bp[jr]);
                inwin = !(d > a.window_bp);
            }
            if (inwin) {
                const double mu_i = a.mean[i];
                const double sum_i = (double)a.sum[i];
                double dot;
                if (MISS) {
                    const double nm_i = dn - (double)a.nobs[i];
                    dot = (double)gg[e] - mu_j * (sum_i - (double)gq[e]) - mu_i * (sum_j - (double)qg[e]) +
                          mu_i * mu_j * (dn - nm_i - nm_j + (double)qq[e]);
                } else {
                    dot = (double)gg[e] - mu_j * sum_i - mu_i * sum_j + mu_i * mu_j * dn;
                }
                double r;
                if (COV) {
                    // projected panel: cross product of the projected columns, minus their
                    // means, times the restandardizing scales
                    const double* ai = a.cov_a + ir * a.cov_k;
                    const double* bj = a.cov_b + jr * a.cov_k;
                    double corr = 0.0;
                    for (int k = 0; k < a.cov_k; ++k) corr = fma(ai[k], bj[k], corr);
                    const double yy = a.scale[i] * s_j * dot + corr - dn * a.ybar[ir] * a.ybar[jr];
                    r = yy * a.tsc[ir] * a.tsc[jr] / dn;
                } else {
                    r = a.scale[i] * s_j * dot / dn;
                }
                const double r2 = r * r;
                v = r2 - (1.0 - r2) / (a.corr_dim - 1.0);
                // one-sided sums: in a diagonal tile the rows take only the pairs on their side
                if (ti == tj && a.side != 0 && (a.side == 1 ? j >= i : j <= i)) v = 0.0;
            }
        }
        r2t[wave][il][lr] = v;
    }
    __syncthreads();
    // rows: L_i += sum_j r2[i][j] a_j ; columns (off-diagonal tiles): L_j += sum_i r2[i][j] a_i.
    // In an off-diagonal tile every row SNP precedes every column SNP: the row sums are
    // right-side sums, the column sums left-side sums.
    const bool diag = ti == tj;
    const bool rows_on = diag || a.side != 1;
    const bool cols_on = !diag && a.side != 2;
    if (half == 0 && rows_on) {
        const int64_t i = i0 + lr;
        if (i < ce) {
            for (int b = 0; b < a.n_bins; ++b) {
                double acc = 0.0;
                for (int c = 0; c < 32; ++c) {
                    const int64_t jc = j0 + c;
                    if (jc < ce) acc += r2t[wave][lr][c] * a.annot[(jc - a.row0) * a.n_bins + b];
                }
                atomicAdd(&a.out[(i - a.row0) * a.n_bins + b], acc);
            }
        }
    } else if (half == 1 && cols_on) {
        const int64_t jc = j0 + lr;
        if (jc < ce) {
            for (int b = 0; b < a.n_bins; ++b) {
                double acc = 0.0;
                for (int rr = 0; rr < 32; ++rr) {
                    const int64_t irr = i0 + rr;
                    if (irr < ce) acc += r2t[wave][rr][lr] * a.annot[(irr - a.row0) * a.n_bins + b];
                }
                atomicAdd(&a.out[(jc - a.row0) * a.n_bins + b], acc);
            }
        }
    }
}

void launch_gram_band(const GramArgs& a, int64_t n_pairs, bool miss, hipStream_t st) {
    if (n_pairs <= 0) return;
    for (int64_t p0 = 0; p0 < n_pairs; p0 += 1 << 30) {
        GramArgs b = a;
        b.pairs = a.pairs + 3 * p0;
        const int64_t np = (n_pairs - p0) < (1 << 30) ? (n_pairs - p0) : (1 << 30);
        const bool cov = a.cov_k > 0;
        if (miss && cov)
            hipLaunchKernelGGL((k_gram_band<true, true>), dim3((unsigned)np), dim3(256), 0, st, b);
        else if (miss)
            hipLaunchKernelGGL((k_gram_band<true, false>), dim3((unsigned)np), dim3(256), 0, st, b);
        else if (cov)
            hipLaunchKernelGGL((k_gram_band<false, true>), dim3((unsigned)np), dim3(256), 0, st, b);
        else
            hipLaunchKernelGGL((k_gram_band<false, false>), dim3((unsigned)np), dim3(256), 0, st, b);
        GPE_HIP(hipGetLastError());
    }
}

// G = C^T C and C^T 1 for the covariate projection of windowed panels: one
// workgroup per (k, q) entry with q >= k (the k == q workgroup also forms the sum).
__global__ __launch_bounds__(256) void k_cov_gram(const double* C, int64_t ldc, int64_t n, int p, double* G,
                                                  double* csum) {
    __shared__ double red[2][4];
    const int k = blockIdx.x, q = blockIdx.y;
    if (q < k) return;
    const int tid = threadIdx.x;
    const double* ck = C + (int64_t)k * ldc;
    const double* cq = C + (int64_t)q * ldc;
    double acc = 0.0, sum = 0.0;
    for (int64_t i = tid; i < n; i += 256) {
        acc = fma(ck[i], cq[i], acc);
        sum += ck[i];
    }
    acc = wave_sum_f64(acc);
    sum = wave_sum_f64(sum);
    if ((tid & 63) == 0) {
        red[0][tid >> 6] = acc;
        red[1][tid >> 6] = sum;
    }
    __syncthreads();
    if (tid == 0) {
        const double g = (red[0][0] + red[0][1]) + (red[0][2] + red[0][3]);
        G[k * p + q] = g;
        G[q * p + k] = g;
        if (q == k) csum[k] = (red[1][0] + red[1][1]) + (red[1][2] + red[1][3]);
    }
}
void launch_cov_gram(const double* C, int64_t ldc, int64_t n, int p, double* G, double* csum, hipStream_t st) {
    if (p <= 0) return;
    hipLaunchKernelGGL(k_cov_gram, dim3((unsigned)p, (unsigned)p), dim3(256), 0, st, C, ldc, n, p, G, csum);
    GPE_HIP(hipGetLastError());
}

__global__ void k_win_cov_prep(const double* u, const double* v, int p, const double* G, const double* csum,
                               const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* inv_sd,
                               int completed, int64_t s, int64_t e, int64_t n, double* A, double* B, double* ybar,
                               double* tsc) {
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t j = s + r;
    if (j >= e) return;
    const double* uj = u + r * p;
    const double* vj = v + r * p;
    double* Aj = A + r * 2 * p;
    double* Bj = B + r * 2 * p;
    // ||x_j||^2 = inv_sd^2 M2 with M2 the centred sum of squares of the (observed) counts
    const double cnt = completed ? (double)n : (double)nobs[j];
    double M2 = 0.0;
    if (cnt > 0) {
        const double sm = (double)sum[j];
        M2 = (double)sumsq[j] - sm * sm / cnt;
        if (M2 < 0.0) M2 = 0.0;
    }
    const double isd = inv_sd[j];
    double yy = isd * isd * M2, cv = 0.0;
    for (int k = 0; k < p; ++k) {
        double gv = 0.0;
        for (int q = 0; q < p; ++q) gv = fma(G[k * p + q], vj[q], gv);
        yy += vj[k] * (gv - 2.0 * uj[k]);
        cv = fma(csum[k], vj[k], cv);
        Aj[k] = gv - uj[k];
        Aj[p + k] = -vj[k];
        Bj[k] = vj[k];
        Bj[p + k] = uj[k];
    }
    // the column sum of x_j is 0 (centred on its own mean), so that of y_j is -(C^T 1) . v_j
    const double ys = -cv;
    const double dn = (double)n;
    ybar[r] = ys / dn;
    double ss = yy - ys * ys / dn;
    if (ss < 0.0 && ss > -1e-12) ss = 0.0;
    double t = 1.0;
    if (ss > 0.0 && isfinite(ss)) {
        const double var = ss / dn;
        if (var > 0.0 && isfinite(var)) t = 1.0 / sqrt(var);
    }
    tsc[r] = t;
}
void launch_win_cov_prep(const double* u, const double* v, int p, const double* G, const double* csum,
                         const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, const double* inv_sd,
                         int completed, int64_t s, int64_t e, int64_t n, double* A, double* B, double* ybar,
                         double* tsc, hipStream_t st) {
    if (e <= s) return;
    hipLaunchKernelGGL(k_win_cov_prep, dim3((unsigned)ceil_div(e - s, 128)), dim3(128), 0, st, u, v, p, G, csum,
                       nobs, sum, sumsq, inv_sd, completed, s, e, n, A, B, ybar, tsc);
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// GxE probe streams: NumPy Philox4x64-10 (counter-based; key per (SNP block,
// probe id) supplied by the host from NumPy's own seeding), consumed as
// Generator.integers(0, 2, int8): uint32 words are the low then high halves of
// each 64-bit output, bytes little-endian, value = high bit of the byte
// (src/summit/ldscore/gwe_ldscore.py:2484-2505, restated natively at
// src/native/gxeldcore.cpp:10002-10138).  Thread per (stream, Philox block):
// 32 probe values.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mulhilo64(uint64_t a, uint64_t b, uint64_t& hi, uint64_t& lo) {
    lo = a * b;
    hi = __umul64hi(a, b);
}

__global__ void k_philox_rademacher(const uint64_t* keys, const int64_t* starts, const int64_t* lens, int n_streams,
                                    int v_count, int64_t row0, int8_t* z, int64_t ldz, int64_t max_blocks) {
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stream = gid / max_blocks;
    const int64_t t = gid % max_blocks;
    if (stream >= n_streams) return;
    const int blk = (int)(stream / v_count);
    const int v = (int)(stream % v_count);
    const int64_t L = lens[blk];
    if (t * 32 >= L) return;
    uint64_t c0 = (uint64_t)t + 1, c1 = 0, c2 = 0, c3 = 0;
    uint64_t k0 = keys[2 * stream], k1 = keys[2 * stream + 1];
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ULL, c0, hi0, lo0);
        mulhilo64(0xCA5A826395121157ULL, c2, hi1, lo1);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0;
        c1 = n1;
        c2 = n2;
        c3 = n3;
        k0 += 0x9E3779B97F4A7C15ULL;
        k1 += 0xBB67AE8584CAA73BULL;
    }
    const uint64_t words[4] = {c0, c1, c2, c3};
    const int64_t base = starts[blk] - row0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
#pragma unroll
        for (int bt = 0; bt < 8; ++bt) {
            const int64_t r = t * 32 + q * 8 + bt;
            if (r < L) {
                const uint32_t byte = (uint32_t)(words[q] >> (8 * bt)) & 0xFFu;
                z[(base + r) * ldz + v] = (byte & 0x80u) ? (int8_t)1 : (int8_t)-1;
            }
        }
    }
}
void launch_philox_rademacher(const uint64_t* keys, const int64_t* starts, const int64_t* lens, int n_blocks,
                              int v_count, int64_t max_len, int64_t row0, int8_t* z, int64_t ldz, hipStream_t st) {
    const int n_streams = n_blocks * v_count;
    const int64_t max_blocks = ceil_div(max_len, 32);
    const int64_t total = (int64_t)n_streams * max_blocks;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_philox_rademacher, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, keys, starts, lens,
                       n_streams, v_count, row0, z, ldz, max_blocks);
    GPE_HIP(hipGetLastError());
}

// out[e][j - r0] = sum_n W[n][e] x_jn^2 for e < n_w (n_w <= EC) and the
// standardized working codes (x = (g - mu) inv_sd, 0 where missing).  One lane
// per SNP; each lane reads 64 codes of its row at a time and, per sample, picks
// x^2 from its four-entry table and adds EC weighted terms.  W is sample-major
// (n_pad, EC), zero beyond n and beyond n_w, so the padded code bytes of a
// row (code 00) add nothing; a sample's EC weights are the same for every lane
// and arrive by scalar loads.
template <int EC>
__global__ __launch_bounds__(256) void k_weighted_sumsq(const uint8_t* __restrict__ snp, int64_t stride, int64_t r0,
                                                        int64_t r1, int64_t n_bytes, const double* __restrict__ mean,
                                                        const double* __restrict__ inv_sd,
                                                        const double* __restrict__ W, int n_w,
                                                        double* __restrict__ out, int64_t ldo) {
    const int64_t j = r0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool live = j < r1;
    const int64_t jj = live ? j : r0;  // idle lanes read a valid row
    const double mu = mean[jj];
    const double x2[4] = {mu * mu, 0.0, (1.0 - mu) * (1.0 - mu), (2.0 - mu) * (2.0 - mu)};
    double acc[EC];
#pragma unroll
    for (int e = 0; e < EC; ++e) acc[e] = 0.0;
    const uint8_t* row = snp + jj * stride;
    for (int64_t b0 = 0; b0 < n_bytes; b0 += 16) {
        const v4u wv = *reinterpret_cast<const v4u*>(row + b0);
        const double* wr = W + b0 * 4 * EC;  // weights of sample 4 * b0
#pragma unroll
        for (int d = 0; d < 4; ++d) {
#pragma unroll 4
            for (int t = 0; t < 16; ++t) {
                const uint32_t c = (wv[d] >> (2 * t)) & 3u;
                const double v = c == 0 ? x2[0] : (c == 2 ? x2[2] : (c == 3 ? x2[3] : 0.0));
                const double* ws = wr + (d * 16 + t) * EC;
#pragma unroll
                for (int e = 0; e < EC; ++e) acc[e] = fma(ws[e], v, acc[e]);
            }
        }
    }
    if (!live) return;
    const double s2 = inv_sd[j] * inv_sd[j];
    for (int e = 0; e < n_w; ++e) out[(int64_t)e * ldo + (j - r0)] = acc[e] * s2;
}
__global__ void k_wss_finalize(const int64_t* __restrict__ ge, const int64_t* __restrict__ me,
                               const int64_t* __restrict__ g2, int64_t ld, int64_t r0, int64_t r1, int n_w,
                               const double* __restrict__ mean, const double* __restrict__ inv_sd,
                               const double* __restrict__ col_scale, const int64_t* __restrict__ col_sum,
                               double* __restrict__ out, int64_t ldo) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t L = r1 - r0;
    if (idx >= L * n_w) return;
    const int e = (int)(idx / L);
    const int64_t jr = idx - (int64_t)e * L;
    const int64_t j = r0 + jr;
    const double mu = mean[j], isd = inv_sd[j];
    const int64_t o = jr * ld + e;
    const double obs = (double)(col_sum[e] - (me ? me[o] : 0));
    const double v = mu * mu * obs - 2.0 * mu * (double)ge[o] + (double)g2[o];
    out[(int64_t)e * ldo + jr] = isd * isd * col_scale[e] * v;
}

void launch_wss_finalize(const int64_t* ge, const int64_t* me, const int64_t* g2, int64_t ld, int64_t r0, int64_t r1,
                         int n_w, const double* mean, const double* inv_sd, const double* col_scale,
                         const int64_t* col_sum, double* out, int64_t ldo, hipStream_t st) {
    const int64_t total = (r1 - r0) * (int64_t)n_w;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_wss_finalize, dim3((unsigned)ceil_div(total, 256)), dim3(256), 0, st, ge, me, g2, ld, r0, r1,
                       n_w, mean, inv_sd, col_scale, col_sum, out, ldo);
    GPE_HIP(hipGetLastError());
}

void launch_weighted_sumsq(const uint8_t* snp, int64_t stride, int64_t r0, int64_t r1, int64_t n, const double* mean,
                           const double* inv_sd, const double* W_sample_major, int ec, int n_w, double* out,
                           int64_t ldo, hipStream_t st) {
    if (r1 <= r0 || n_w <= 0) return;
    const int64_t n_bytes = round_up(ceil_div(n, 4), 16);
    const dim3 grid((unsigned)ceil_div(r1 - r0, 256));
    if (ec == 8)
        hipLaunchKernelGGL(k_weighted_sumsq<8>, grid, dim3(256), 0, st, snp, stride, r0, r1, n_bytes, mean, inv_sd,
                           W_sample_major, n_w, out, ldo);
    else if (ec == 16)
        hipLaunchKernelGGL(k_weighted_sumsq<16>, grid, dim3(256), 0, st, snp, stride, r0, r1, n_bytes, mean, inv_sd,
                           W_sample_major, n_w, out, ldo);
    else
        throw Error("weighted sums of squares take 8 or 16 weights per pass");
    GPE_HIP(hipGetLastError());
}

// s_i = sqrt(N / M2_i) (1 when M2_i == 0), M2 over the observed (mean) or completed calls
__global__ void k_window_scale(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                               int completed, double* scale) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    const double cnt = completed ? (double)n : (double)nobs[j];
    double M2 = 0.0;
    if (cnt > 0) {
        const double s = (double)sum[j];
        M2 = (double)sumsq[j] - s * s / cnt;
        if (M2 < 0.0 && M2 > -1e-12) M2 = 0.0;
    }
    scale[j] = (M2 > 0.0 && isfinite(M2)) ? sqrt((double)n / M2) : 1.0;
}
void launch_window_scale(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t m, int64_t n,
                         int completed, double* scale, hipStream_t st) {
    if (m <= 0) return;
    hipLaunchKernelGGL(k_window_scale, dim3((unsigned)ceil_div(m, 256)), dim3(256), 0, st, nobs, sum, sumsq, m, n,
                       completed, scale);
    GPE_HIP(hipGetLastError());
}

// ---------------------------------------------------------------------------
// Synthetic genotypes written straight into the SNP-major layout.
// Two haplotypes per sample follow a per-block Markov chain along SNPs: at a
// block start each allele is a fresh Bernoulli(MAF_j) draw, inside a block it
// copies the previous SNP's allele with probability copy_prob, else it is
// redrawn (copy_prob = 0 gives i.i.d. Binomial(2, MAF_j) genotypes).  A call
// is missing with probability miss_rate.  MAF_j = lo + (hi-lo) * U_j.
// One thread owns 4 samples (one byte column) x one SNP block and draws from
// a xorshift64* stream keyed by (seed, block, byte column).
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t xs64(uint64_t& s) {
    s ^= s >> 12;
    s ^= s << 25;
    s ^= s >> 27;
    return s * 0x2545F4914F6CDD1DULL;
}

__global__ void k_synth_markov(uint8_t* snp, int64_t stride, int64_t m, int64_t n, uint64_t seed, double maf_lo,
                               double maf_hi, uint64_t miss32, int block_len, uint32_t copy16) {
    const int64_t bcol = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t blk = blockIdx.y + (int64_t)blockIdx.z * 65535;
    const int64_t nbytes = ceil_div(n, 4);
    if (bcol >= nbytes) return;
    const int64_t j0 = blk * block_len;
    if (j0 >= m) return;
    uint64_t st = mix64(seed ^ mix64((uint64_t)blk * 0x9E3779B97F4A7C15ULL ^ mix64((uint64_t)bcol + 0x632BE59BD9B4E019ULL)));
    if (st == 0) st = 0x9E3779B97F4A7C15ULL;
    uint32_t a0 = 0, a1 = 0;  // bit t: allele of haplotype 0/1 of sample 4*bcol+t
    const int64_t jend = (j0 + block_len < m) ? j0 + block_len : m;
    const uint64_t maf_key = mix64(seed + 0xD1B54A32D192ED03ULL);
    for (int64_t j = j0; j < jend; ++j) {
        const uint64_t hj = mix64(maf_key ^ (uint64_t)j);
        const double u = (double)(hj >> 11) * (1.0 / 9007199254740992.0);
        const double p = maf_lo + (maf_hi - maf_lo) * u;
        const uint32_t p16 = (uint32_t)(p * 65536.0);
        uint32_t byte = 0;
        for (int t = 0; t < 4; ++t) {
            const uint64_t r1 = xs64(st);
            const uint64_t r2 = xs64(st);
            const uint32_t f0 = ((uint32_t)(r1 & 0xFFFF) < p16) ? 1u : 0u;
            const uint32_t f1 = ((uint32_t)((r1 >> 16) & 0xFFFF) < p16) ? 1u : 0u;
            const bool c0 = (j > j0) && ((uint32_t)((r1 >> 32) & 0xFFFF) < copy16);
            const bool c1 = (j > j0) && ((uint32_t)(r1 >> 48) < copy16);
            const uint32_t h0 = c0 ? ((a0 >> t) & 1u) : f0;
            const uint32_t h1 = c1 ? ((a1 >> t) & 1u) : f1;
            a0 = (a0 & ~(1u << t)) | (h0 << t);
            a1 = (a1 & ~(1u << t)) | (h1 << t);
            const uint32_t g = h0 + h1;
            uint32_t code = g == 0 ? 0u : (g == 1 ? 2u : 3u);
            if ((r2 >> 32) < miss32) code = 1u;
            if (bcol * 4 + t >= n) code = 0u;
            byte |= code << (2 * t);
        }
        snp[j * stride + bcol] = (uint8_t)byte;
    }
}
void launch_synth_markov(uint8_t* snp, int64_t stride, int64_t m, int64_t n, uint64_t seed, double maf_lo,
                         double maf_hi, double miss_rate, int block_len, double copy_prob, int iid,
                         hipStream_t st) {
    if (m <= 0 || n <= 0) return;
    const int64_t nblk = ceil_div(m, block_len);
    const uint64_t miss32 = (uint64_t)(miss_rate * 4294967296.0);
    const uint32_t copy16 = iid ? 0u : (uint32_t)(copy_prob * 65536.0);
    dim3 grid((unsigned)ceil_div(ceil_div(n, 4), 256), (unsigned)(nblk < 65535 ? nblk : 65535),
              (unsigned)ceil_div(nblk, 65535));
    hipLaunchKernelGGL(k_synth_markov, grid, dim3(256), 0, st, snp, stride, m, n, seed, maf_lo, maf_hi, miss32,
                       block_len, copy16);
    GPE_HIP(hipGetLastError());
}


// ---------------------------------------------------------------------------
// Small-rank FP64 projection of columns: x_q <- x_q - A (B x_q), optionally
// followed by centring (x_q <- x_q - mean(x_q)).  A is n x p column-major
// (A[i + k lda]), B is p x n row-major (B[k ldb + i]), x column q at
// x[q ldx + i].  The covariate projection I - C R of phase 1/2 inputs and
// the GxE projection I - Q Q^T plus intercept.  All reductions run in a
// fixed order.
// ---------------------------------------------------------------------------
constexpr int kProjRows = 32;  // rows of B per coefficient pass

// T[k][q] = sum_i B[k0 + k][i] x[q][i] (k < kb), S[q] = sum_i x[q][i] (when S != null)
__global__ __launch_bounds__(256) void k_proj_coef(const double* __restrict__ x, int64_t ldx, int64_t n,
                                                   const double* __restrict__ B, int64_t ldb, int k0, int kb,
                                                   double* __restrict__ T, int64_t ldt, double* __restrict__ S) {
    __shared__ double red[kProjRows + 1][4];
    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    double acc[kProjRows + 1];
#pragma unroll
    for (int k = 0; k <= kProjRows; ++k) acc[k] = 0.0;
    const double* xq = x + (int64_t)q * ldx;
    for (int64_t i = tid; i < n; i += 256) {
        const double v = xq[i];
        acc[kProjRows] += v;
#pragma unroll
        for (int k = 0; k < kProjRows; ++k)
            if (k < kb) acc[k] = fma(B[(int64_t)(k0 + k) * ldb + i], v, acc[k]);
    }
    const int wave = tid >> 6, lane = tid & 63;
#pragma unroll
    for (int k = 0; k <= kProjRows; ++k) {
        const double w = wave_sum_f64(acc[k]);
        if (lane == 0) red[k][wave] = w;
    }
    __syncthreads();
    if (tid <= kProjRows) {
        const double v = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
        if (tid < kb) T[(int64_t)(k0 + tid) * ldt + q] = v;
        if (tid == kProjRows && S) S[q] = v;
    }
}

// a[k] = sum_i A[i + k lda]
__global__ __launch_bounds__(256) void k_proj_colsum(const double* __restrict__ A, int64_t lda, int64_t n,
                                                     double* __restrict__ a) {
    __shared__ double red[256];
    const int k = blockIdx.x, tid = threadIdx.x;
    double s = 0.0;
    for (int64_t i = tid; i < n; i += 256) s += A[i + (int64_t)k * lda];
    red[tid] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) red[tid] += red[tid + w];
        __syncthreads();
    }
    if (tid == 0) a[k] = red[0];
}

// mean[q] = (S[q] - sum_k a[k] T[k][q]) / n: the mean of the projected column
__global__ void k_proj_mean(const double* __restrict__ S, const double* __restrict__ a, const double* __restrict__ T,
                            int64_t ldt, int p, int q_count, int64_t n, double* __restrict__ mean) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= q_count) return;
    double v = S[q];
    for (int k = 0; k < p; ++k) v -= a[k] * T[(int64_t)k * ldt + q];
    mean[q] = v / (double)n;
}

// x[q][i] -= sum_k A[i][k] T[k][q] (+ mean[q])
__global__ __launch_bounds__(256) void k_proj_apply(double* __restrict__ x, int64_t ldx, int64_t n,
                                                    const double* __restrict__ A, int64_t lda,
                                                    const double* __restrict__ T, int64_t ldt, int p,
                                                    const double* __restrict__ mean) {
    const int q = blockIdx.y;
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    double v = x[(int64_t)q * ldx + i];
    for (int k = 0; k < p; ++k) v -= A[i + (int64_t)k * lda] * T[(int64_t)k * ldt + q];
    if (mean) v -= mean[q];
    x[(int64_t)q * ldx + i] = v;
}

// s2[q] = sum_i x[q][i]^2 (fixed-order reduction)
__global__ __launch_bounds__(256) void k_col_sumsq(const double* __restrict__ x, int64_t ldx, int64_t n,
                                                   double* __restrict__ s2) {
    __shared__ double red[4];
    const int q = blockIdx.x, tid = threadIdx.x;
    const double* xq = x + (int64_t)q * ldx;
    double acc = 0.0;
    for (int64_t i = tid; i < n; i += 256) acc = fma(xq[i], xq[i], acc);
    const double w = wave_sum_f64(acc);
    if ((tid & 63) == 0) red[tid >> 6] = w;
    __syncthreads();
    if (tid == 0) s2[q] = (red[0] + red[1]) + (red[2] + red[3]);
}

void launch_basis_coef(const double* x, int64_t ldx, int64_t n, int q_count, const double* B, int64_t ldb, int p,
                       double* T, int64_t ldt, double* s2, hipStream_t st) {
    if (n <= 0 || q_count <= 0) return;
    for (int k0 = 0; k0 < p; k0 += kProjRows) {
        const int kb = std::min(kProjRows, p - k0);
        hipLaunchKernelGGL(k_proj_coef, dim3((unsigned)q_count), dim3(256), 0, st, x, ldx, n, B, ldb, k0, kb, T, ldt,
                           nullptr);
        GPE_HIP(hipGetLastError());
    }
    if (s2) {
        hipLaunchKernelGGL(k_col_sumsq, dim3((unsigned)q_count), dim3(256), 0, st, x, ldx, n, s2);
        GPE_HIP(hipGetLastError());
    }
}

void launch_col_sums(const double* x, int64_t ldx, int64_t n, int q_count, double* S, hipStream_t st) {
    if (q_count <= 0) return;
    hipLaunchKernelGGL(k_proj_coef, dim3((unsigned)q_count), dim3(256), 0, st, x, ldx, n, nullptr, (int64_t)0, 0, 0,
                       nullptr, (int64_t)q_count, S);
    GPE_HIP(hipGetLastError());
}

void launch_project(double* x, int64_t ldx, int64_t n, int q_count, const double* A, int64_t lda, const double* B,
                    int64_t ldb, int p, bool center, double* T, double* S, double* a, double* mean, hipStream_t st) {
    if (n <= 0 || q_count <= 0) return;
    for (int k0 = 0; k0 < p || (k0 == 0 && center); k0 += kProjRows) {
        const int kb = std::max(0, std::min(kProjRows, p - k0));
        hipLaunchKernelGGL(k_proj_coef, dim3((unsigned)q_count), dim3(256), 0, st, x, ldx, n, B, ldb, k0, kb, T,
                           (int64_t)q_count, (k0 == 0 && center) ? S : nullptr);
        GPE_HIP(hipGetLastError());
        if (p == 0) break;
    }
    if (center) {
        if (p > 0) {
            hipLaunchKernelGGL(k_proj_colsum, dim3((unsigned)p), dim3(256), 0, st, A, lda, n, a);
            GPE_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(k_proj_mean, dim3((unsigned)ceil_div(q_count, 256)), dim3(256), 0, st, S, a, T,
                           (int64_t)q_count, p, q_count, n, mean);
        GPE_HIP(hipGetLastError());
    }
    if (p == 0 && !center) return;
    hipLaunchKernelGGL(k_proj_apply, dim3((unsigned)ceil_div(n, 256), (unsigned)q_count), dim3(256), 0, st, x, ldx, n,
                       A, lda, T, (int64_t)q_count, p, center ? mean : nullptr);
    GPE_HIP(hipGetLastError());
}

}  // namespace gpe
