// Learned LD-score predictor: a three-layer MLP (in -> 128 -> 128 -> 1, GELU)
// trained and evaluated by hand-written gfx950 kernels.  Weight matrices are
// FP32 masters used as BF16; inputs and hidden activations enter each product
// as BF16 and every product accumulates in FP32 on v_mfma_f32_32x32x16_bf16.
// Biases stay FP32.  Adam keeps FP32 moments.  No reference counterpart: the
// network is the control variate of the stochastic LD-score estimator.
//
// Row-major activations [rows][128]; the weight-gradient products contract
// over rows, so their operands are also kept transposed ([128][ldT], rows
// contiguous).  Every reduction over rows is two-level (per-workgroup
// partials, then a fixed-order sum), so results do not depend on scheduling.
// Entry points only enqueue work on the caller's stream and never allocate
// after gpe_mlp_create, so a training step can be captured in a HIP graph.
#include "../../include/summit_gpe.h"
#include "gpe_internal.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using namespace gpe;

namespace gpe_mlpk {

constexpr int kH = 128;        // hidden width
constexpr int kInPad = 32;     // input width after zero padding
constexpr int kRowTile = 64;   // rows per workgroup of the row kernels
constexpr int kWChunk = 512;   // rows per partial of the weight-gradient products
constexpr int64_t kNParams = (int64_t)kH * kInPad + kH + (int64_t)kH * kH + kH + kH + 1;
// offsets of the pieces in the flat parameter vector
constexpr int64_t kOffW1 = 0;
constexpr int64_t kOffB1 = kOffW1 + (int64_t)kH * kInPad;
constexpr int64_t kOffW2 = kOffB1 + kH;
constexpr int64_t kOffB2 = kOffW2 + (int64_t)kH * kH;
constexpr int64_t kOffW3 = kOffB2 + kH;
constexpr int64_t kOffB3 = kOffW3 + kH;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t bf16_bits(float f) {
    // round to nearest, ties to even (inputs are finite)
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf16_float(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ float gelu_f(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float x) {
    return 0.5f * (1.0f + erff(x * 0.70710678118654752f)) + x * 0.39894228040143268f * expf(-0.5f * x * x);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) { return *reinterpret_cast<const bf16x8*>(p); }

// Eight consecutive bf16 values starting at p, zero at and beyond `valid` elements.
__device__ __forceinline__ bf16x8 load8_masked(const uint16_t* p, int64_t valid) {
    if (valid >= 8) return load8(p);
    uint16_t t[8];
#pragma unroll
    for (int e = 0; e < 8; ++e) t[e] = (e < valid) ? p[e] : (uint16_t)0;
    bf16x8 r;
    __builtin_memcpy(&r, t, 16);
    return r;
}

__device__ __forceinline__ f32x16 mfma_bf16(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Row-tile product: out[r][c] = sum_k A[r][k] * B[c][k] for a 64-row tile and
// all 128 output columns.  Wave w owns rows 32 (w >> 1) .. +32 and columns
// 64 (w & 1) .. +64 (two 32x32 MFMA tiles).  Lane (half, lr) supplies A row
// lr and B row lr of each tile, eight consecutive k from 8 half.  Accumulator
// element e of tile t is row 8 (e >> 2) + 4 half + (e & 3), column 32 t + lr.
// ---------------------------------------------------------------------------
template <int KD>
__device__ __forceinline__ void row_tile_product(const uint16_t* __restrict__ A, int64_t lda, int64_t rows,
                                                 const uint16_t* __restrict__ B, int64_t r0, int c0, int half,
                                                 int lr, f32x16& acc0, f32x16& acc1) {
    const int64_t ra = r0 + lr;
    const bool rv = ra < rows;
    const uint16_t* pa = A + (rv ? ra : 0) * lda + half * 8;
    const uint16_t* pb0 = B + (int64_t)(c0 + lr) * KD + half * 8;
    const uint16_t* pb1 = B + (int64_t)(c0 + 32 + lr) * KD + half * 8;
    const bf16x8 zero = bf16x8{};
#pragma unroll
    for (int s = 0; s < KD / 16; ++s) {
        const bf16x8 a = rv ? load8(pa + s * 16) : zero;
        acc0 = mfma_bf16(a, load8(pb0 + s * 16), acc0);
        acc1 = mfma_bf16(a, load8(pb1 + s * 16), acc1);
    }
}

// Layer forward: z = A W^T + b, h = bf16(GELU(z)); optionally z (FP32,
// row-major) and h^T (bf16, [128][ldT]) for the backward pass.
template <int KD>
__global__ __launch_bounds__(256) void k_mlp_fwd(const uint16_t* __restrict__ x, int64_t ldx, int64_t rows,
                                                 const uint16_t* __restrict__ W, const float* __restrict__ bias,
                                                 float* __restrict__ z, uint16_t* __restrict__ h,
                                                 uint16_t* __restrict__ hT, int64_t ldT) {
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, lr = lane & 31;
    const int64_t r0 = (int64_t)blockIdx.x * kRowTile + (wave >> 1) * 32;
    const int c0 = (wave & 1) * 64;
    f32x16 acc[2] = {f32x16{}, f32x16{}};
    row_tile_product<KD>(x, ldx, rows, W, r0, c0, half, lr, acc[0], acc[1]);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = c0 + 32 * t + lr;
        const float bc = bias[col];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int64_t rb = r0 + 8 * g + 4 * half;
            uint16_t hv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const float zz = acc[t][4 * g + q] + bc;
                hv[q] = bf16_bits(gelu_f(zz));
                const int64_t r = rb + q;
                if (r < rows) {
                    if (z) z[r * kH + col] = zz;
                    h[r * kH + col] = hv[q];
                }
            }
            if (hT) {
                // rb < rows rounded up to the row tile <= ldT; rows >= `rows` are never read back
                const uint2 pk = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16),
                                            (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
                *reinterpret_cast<uint2*>(hT + (int64_t)col * ldT + rb) = pk;
            }
        }
    }
}

// Output head: y = w3 . h2 + b3; with a target, dy = 2 (y - t) / rows and the
// workgroup's sum of (y - t)^2.  One thread per row, 256 rows per workgroup.
__global__ __launch_bounds__(256) void k_mlp_head(const uint16_t* __restrict__ h2, int64_t rows,
                                                  const uint16_t* __restrict__ w3b, const float* __restrict__ b3,
                                                  const float* __restrict__ target, float inv_rows,
                                                  float* __restrict__ y, float* __restrict__ dy,
                                                  float* __restrict__ loss_part) {
    __shared__ float w[kH];
    __shared__ float red[256];
    const int tid = threadIdx.x;
    if (tid < kH) w[tid] = bf16_float(w3b[tid]);
    __syncthreads();
    const int64_t r = (int64_t)blockIdx.x * 256 + tid;
    float sq = 0.0f;
    if (r < rows) {
        const uint16_t* p = h2 + r * kH;
        float acc = 0.0f;
#pragma unroll
        for (int c8 = 0; c8 < kH / 8; ++c8) {
            const uint4 v = *reinterpret_cast<const uint4*>(p + 8 * c8);
            const uint32_t u[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                acc = fmaf(bf16_float((uint16_t)(u[q] & 0xFFFFu)), w[8 * c8 + 2 * q], acc);
                acc = fmaf(bf16_float((uint16_t)(u[q] >> 16)), w[8 * c8 + 2 * q + 1], acc);
            }
        }
        const float yy = acc + b3[0];
        y[r] = yy;
        if (target) {
            const float d = yy - target[r];
            dy[r] = 2.0f * d * inv_rows;
            sq = d * d;
        }
    }
    if (!target) return;
    red[tid] = sq;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (tid < s) red[tid] += red[tid + s];
        __syncthreads();
    }
    if (tid == 0) loss_part[blockIdx.x] = red[0];
}

// Head backward: dz2 = dy w3 GELU'(z2) (bf16 row-major and transposed), and
// per-workgroup partial sums for db2 (from FP32 dz2), dw3 = sum dy h2, db3.
// Thread (rh, c): column c, rows rh*32 .. +32 of the 64-row tile.
__global__ __launch_bounds__(256) void k_mlp_head_bwd(const float* __restrict__ dy, const float* __restrict__ z2,
                                                      const uint16_t* __restrict__ h2, int64_t rows,
                                                      const uint16_t* __restrict__ w3b, uint16_t* __restrict__ dz2,
                                                      uint16_t* __restrict__ dz2T, int64_t ldT,
                                                      float* __restrict__ p_b2, float* __restrict__ p_w3,
                                                      float* __restrict__ p_b3) {
    __shared__ float sb[2][kH], sw[2][kH], sd[2];
    const int tid = threadIdx.x, c = tid & (kH - 1), rh = tid >> 7;
    const int64_t rbase = (int64_t)blockIdx.x * kRowTile + rh * 32;
    const float w = bf16_float(w3b[c]);
    float accb = 0.0f, accw = 0.0f, accd = 0.0f;
    uint16_t pk[8];
#pragma unroll 8
    for (int i = 0; i < 32; ++i) {
        const int64_t r = rbase + i;
        float g = 0.0f;
        if (r < rows) {
            const float d = dy[r];
            g = d * w * gelu_grad_f(z2[r * kH + c]);
            accw = fmaf(d, bf16_float(h2[r * kH + c]), accw);
            accd += d;
        }
        const uint16_t gb = bf16_bits(g);
        if (r < rows) dz2[r * kH + c] = gb;
        accb += g;
        pk[i & 7] = gb;
        if ((i & 7) == 7) {
            uint4 v;
            v.x = (uint32_t)pk[0] | ((uint32_t)pk[1] << 16);
            v.y = (uint32_t)pk[2] | ((uint32_t)pk[3] << 16);
            v.z = (uint32_t)pk[4] | ((uint32_t)pk[5] << 16);
            v.w = (uint32_t)pk[6] | ((uint32_t)pk[7] << 16);
            *reinterpret_cast<uint4*>(dz2T + (int64_t)c * ldT + rbase + (i - 7)) = v;
        }
    }
    sb[rh][c] = accb;
    sw[rh][c] = accw;
    if (c == 0) sd[rh] = accd;
    __syncthreads();
    if (rh == 0) {
        p_b2[(int64_t)blockIdx.x * kH + c] = sb[0][c] + sb[1][c];
        p_w3[(int64_t)blockIdx.x * kH + c] = sw[0][c] + sw[1][c];
        if (c == 0) p_b3[blockIdx.x] = sd[0] + sd[1];
    }
}

// dh1 = dz2 W2 (B operand W2^T), dz1 = dh1 GELU'(z1): dz1^T in bf16 and the
// workgroup's column sums of FP32 dz1 (db1 partials).
__global__ __launch_bounds__(256) void k_mlp_dh1(const uint16_t* __restrict__ dz2, int64_t rows,
                                                 const uint16_t* __restrict__ W2T, const float* __restrict__ z1,
                                                 uint16_t* __restrict__ dz1T, int64_t ldT, float* __restrict__ p_b1) {
    __shared__ float cs[4][2][64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, lr = lane & 31;
    const int64_t r0 = (int64_t)blockIdx.x * kRowTile + (wave >> 1) * 32;
    const int c0 = (wave & 1) * 64;
    f32x16 acc[2] = {f32x16{}, f32x16{}};
    row_tile_product<kH>(dz2, kH, rows, W2T, r0, c0, half, lr, acc[0], acc[1]);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const int col = c0 + 32 * t + lr;
        float csum = 0.0f;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int64_t rb = r0 + 8 * g + 4 * half;
            uint16_t hv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int64_t r = rb + q;
                float v = 0.0f;
                if (r < rows) v = acc[t][4 * g + q] * gelu_grad_f(z1[r * kH + col]);
                csum += v;
                hv[q] = bf16_bits(v);
            }
            const uint2 pk = make_uint2((uint32_t)hv[0] | ((uint32_t)hv[1] << 16),
                                        (uint32_t)hv[2] | ((uint32_t)hv[3] << 16));
            *reinterpret_cast<uint2*>(dz1T + (int64_t)col * ldT + rb) = pk;
        }
        cs[wave][half][32 * t + lr] = csum;
    }
    __syncthreads();
    if (tid < kH) {
        const int col = tid, wc = col >> 6, cc = col & 63;  // waves wc (rows 0-31) and wc + 2 (rows 32-63)
        p_b1[(int64_t)blockIdx.x * kH + col] =
            (cs[wc][0][cc] + cs[wc][1][cc]) + (cs[wc + 2][0][cc] + cs[wc + 2][1][cc]);
    }
}

// Weight gradient partial: part[chunk][i][j] = sum_{r in chunk} A[i][r] B[j][r]
// for i < 128, j < NC; A = dz^T [128][lda], B = h^T or x^T [NC][ldb], rows
// contiguous.  Wave w owns rows 32 w .. +32 of the partial and all NC/32
// column tiles; rows at and beyond `rows` are masked to zero.
template <int NC>
__global__ __launch_bounds__(256) void k_mlp_wgrad(const uint16_t* __restrict__ A, int64_t lda,
                                                   const uint16_t* __restrict__ B, int64_t ldb, int64_t rows,
                                                   float* __restrict__ part) {
    constexpr int NT = NC / 32;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63, half = lane >> 5, lr = lane & 31;
    const int64_t kb = (int64_t)blockIdx.x * kWChunk;
    const int64_t ke = (kb + kWChunk < rows) ? kb + kWChunk : rows;
    const uint16_t* pa = A + (int64_t)(32 * wave + lr) * lda;
    f32x16 acc[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = f32x16{};
    for (int64_t k = kb; k < ke; k += 16) {
        const int64_t kk = k + 8 * half;
        const int64_t valid = ke - kk;
        const bf16x8 a = valid > 0 ? load8_masked(pa + kk, valid) : bf16x8{};
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const uint16_t* pb = B + (int64_t)(32 * t + lr) * ldb;
            const bf16x8 b = valid > 0 ? load8_masked(pb + kk, valid) : bf16x8{};
            acc[t] = mfma_bf16(a, b, acc[t]);
        }
    }
    float* out = part + (int64_t)blockIdx.x * kH * NC;
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int i = 32 * wave + 8 * (e >> 2) + 4 * half + (e & 3);
            out[(int64_t)i * NC + 32 * t + lr] = acc[t][e];
        }
}

// Fixed-order sum of per-workgroup loss partials: loss = sum / rows.
__global__ __launch_bounds__(256) void k_mlp_loss(const float* __restrict__ part, int64_t n, float inv_rows,
                                                  float* __restrict__ loss) {
    __shared__ double red[256];
    const int tid = threadIdx.x;
    double s = 0.0;
    for (int64_t i = tid; i < n; i += 256) s += (double)part[i];
    red[tid] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (tid < k) red[tid] += red[tid + k];
        __syncthreads();
    }
    if (tid == 0) loss[0] = (float)(red[0] * (double)inv_rows);
}

struct Piece {
    const float* part;  // partials [n_part][size] (stride = size)
    int64_t n_part;
    int64_t size;
    int64_t offset;
};

struct AdamArgs {
    Piece pc[6];
    float* p;
    float* m;
    float* v;
    float* grad;
    uint16_t* W1b;
    uint16_t* W2b;
    uint16_t* W2Tb;
    uint16_t* w3b;
    int32_t* step;
    float lr, beta1, beta2, eps, wd;
    int update;
};

// Gradient = fixed-order sum of each piece's partials; Adam(W) update of the
// FP32 masters (update != 0) and the bf16 copies the kernels read.
__global__ __launch_bounds__(256) void k_mlp_adam(AdamArgs a) {
    const int64_t idx = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (idx >= kNParams) return;
    int q = 0;
#pragma unroll
    for (int i = 1; i < 6; ++i)
        if (idx >= a.pc[i].offset) q = i;
    const Piece pc = a.pc[q];
    const int64_t j = idx - pc.offset;
    float g = 0.0f;
    for (int64_t c = 0; c < pc.n_part; ++c) g += pc.part[c * pc.size + j];
    a.grad[idx] = g;
    float p = a.p[idx];
    if (a.update) {
        const double t = (double)a.step[0] + 1.0;
        const float m = a.beta1 * a.m[idx] + (1.0f - a.beta1) * g;
        const float v = a.beta2 * a.v[idx] + (1.0f - a.beta2) * g * g;
        a.m[idx] = m;
        a.v[idx] = v;
        const float mh = m / (float)(1.0 - pow((double)a.beta1, t));
        const float vh = v / (float)(1.0 - pow((double)a.beta2, t));
        p = p - a.lr * (mh / (sqrtf(vh) + a.eps) + a.wd * p);
        a.p[idx] = p;
    }
    const uint16_t b = bf16_bits(p);
    if (q == 0) a.W1b[j] = b;
    if (q == 2) {
        a.W2b[j] = b;
        a.W2Tb[(j % kH) * kH + j / kH] = b;
    }
    if (q == 4) a.w3b[j] = b;
}

__global__ void k_mlp_step_inc(int32_t* step, int32_t inc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) step[0] += inc;
}

__global__ void k_mlp_zero_f32(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

template <class F>
int mlp_guarded(F&& f) {
    try {
        f();
        return GPE_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return std::strncmp(e.what(), "HIP error", 9) == 0 ? GPE_EDEVICE : GPE_EINVAL;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GPE_EINTERNAL;
    } catch (...) {
        set_last_error("unknown error");
        return GPE_EINTERNAL;
    }
}

void require(bool ok, const std::string& msg) {
    if (!ok) throw Error(msg);
}

}  // namespace gpe_mlpk

using namespace gpe_mlpk;

struct gpe_mlp {
    int device = 0;
    int32_t in_dim = 0;
    int64_t max_rows = 0, rows_pad = 0, ld_t = 0;
    int64_t n_tiles = 0, n_heads = 0, n_wchunks = 0;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr;
    uint16_t *W1b = nullptr, *W2b = nullptr, *W2Tb = nullptr, *w3b = nullptr;
    int32_t* step = nullptr;
    float *z1 = nullptr, *z2 = nullptr, *y = nullptr, *dy = nullptr, *loss_part = nullptr;
    uint16_t *h1 = nullptr, *h1T = nullptr, *h2 = nullptr, *dz2 = nullptr, *dz2T = nullptr, *dz1T = nullptr;
    float *p_b2 = nullptr, *p_w3 = nullptr, *p_b3 = nullptr, *p_b1 = nullptr, *p_W2 = nullptr, *p_W1 = nullptr;
    std::vector<void*> allocs;

    template <class T>
    T* alloc(int64_t count) {
        void* q = nullptr;
        GPE_HIP(hipMalloc(&q, (size_t)std::max<int64_t>(count, 1) * sizeof(T)));
        allocs.push_back(q);
        return static_cast<T*>(q);
    }
    void release() {
        for (void* q : allocs) (void)hipFree(q);
        allocs.clear();
    }
};

namespace {

hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

void check_x(const gpe_mlp* net, const uint16_t* x, int64_t ldx) {
    require(x != nullptr, "feature matrix is null");
    require(ldx >= kInPad && ldx % 8 == 0, "feature rows must hold >= 32 bf16 values, a multiple of 8");
    require((reinterpret_cast<uintptr_t>(x) & 15) == 0, "feature matrix must be 16-byte aligned");
    (void)net;
}

// Bf16 copies of the masters and the gradient buffer from the current
// parameters (no update).
void refresh_copies(gpe_mlp* net, hipStream_t st) {
    AdamArgs a{};
    const int64_t sizes[6] = {kOffB1 - kOffW1, kOffW2 - kOffB1, kOffB2 - kOffW2, kOffW3 - kOffB2, kOffB3 - kOffW3, 1};
    const int64_t offs[6] = {kOffW1, kOffB1, kOffW2, kOffB2, kOffW3, kOffB3};
    for (int i = 0; i < 6; ++i) a.pc[i] = Piece{nullptr, 0, sizes[i], offs[i]};
    a.p = net->p;
    a.m = net->m;
    a.v = net->v;
    a.grad = net->grad;
    a.W1b = net->W1b;
    a.W2b = net->W2b;
    a.W2Tb = net->W2Tb;
    a.w3b = net->w3b;
    a.step = net->step;
    a.update = 0;
    hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)ceil_div(kNParams, 256)), dim3(256), 0, st, a);
    GPE_HIP(hipGetLastError());
}

// Forward over rows [0, rows) of one chunk that fits the workspace.
void forward_chunk(gpe_mlp* net, const uint16_t* x, int64_t ldx, int64_t rows, bool train, hipStream_t st) {
    const unsigned tiles = (unsigned)ceil_div(rows, kRowTile);
    hipLaunchKernelGGL(k_mlp_fwd<kInPad>, dim3(tiles), dim3(256), 0, st, x, ldx, rows, net->W1b, net->p + kOffB1,
                       train ? net->z1 : nullptr, net->h1, train ? net->h1T : nullptr, net->ld_t);
    GPE_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_mlp_fwd<kH>, dim3(tiles), dim3(256), 0, st, net->h1, (int64_t)kH, rows, net->W2b,
                       net->p + kOffB2, train ? net->z2 : nullptr, net->h2, nullptr, net->ld_t);
    GPE_HIP(hipGetLastError());
}

}  // namespace

extern "C" {

int gpe_mlp_n_params(int64_t* n) {
    return mlp_guarded([&] {
        require(n != nullptr, "n is null");
        *n = kNParams;
    });
}

int gpe_mlp_create(int32_t in_dim, int64_t max_rows, gpe_mlp** out) {
    return mlp_guarded([&] {
        require(out != nullptr, "out is null");
        require(in_dim >= 1 && in_dim <= kInPad, "in_dim must be in [1, 32]");
        require(max_rows >= 1 && max_rows <= (int64_t)1 << 26, "max_rows must be in [1, 2^26]");
        gpe_mlp* net = new gpe_mlp();
        try {
            GPE_HIP(hipGetDevice(&net->device));
            net->in_dim = in_dim;
            net->max_rows = max_rows;
            net->rows_pad = round_up(max_rows, kRowTile);
            net->ld_t = net->rows_pad;
            net->n_tiles = net->rows_pad / kRowTile;
            net->n_heads = ceil_div(net->rows_pad, 256);
            net->n_wchunks = ceil_div(net->rows_pad, kWChunk);
            const int64_t R = net->rows_pad;
            net->p = net->alloc<float>(kNParams);
            net->m = net->alloc<float>(kNParams);
            net->v = net->alloc<float>(kNParams);
            net->grad = net->alloc<float>(kNParams);
            net->W1b = net->alloc<uint16_t>((int64_t)kH * kInPad);
            net->W2b = net->alloc<uint16_t>((int64_t)kH * kH);
            net->W2Tb = net->alloc<uint16_t>((int64_t)kH * kH);
            net->w3b = net->alloc<uint16_t>(kH);
            net->step = net->alloc<int32_t>(1);
            net->z1 = net->alloc<float>(R * kH);
            net->z2 = net->alloc<float>(R * kH);
            net->y = net->alloc<float>(R);
            net->dy = net->alloc<float>(R);
            net->loss_part = net->alloc<float>(net->n_heads);
            net->h1 = net->alloc<uint16_t>(R * kH);
            net->h1T = net->alloc<uint16_t>(R * kH);
            net->h2 = net->alloc<uint16_t>(R * kH);
            net->dz2 = net->alloc<uint16_t>(R * kH);
            net->dz2T = net->alloc<uint16_t>(R * kH);
            net->dz1T = net->alloc<uint16_t>(R * kH);
            net->p_b2 = net->alloc<float>(net->n_tiles * kH);
            net->p_w3 = net->alloc<float>(net->n_tiles * kH);
            net->p_b3 = net->alloc<float>(net->n_tiles);
            net->p_b1 = net->alloc<float>(net->n_tiles * kH);
            net->p_W2 = net->alloc<float>(net->n_wchunks * kH * kH);
            net->p_W1 = net->alloc<float>(net->n_wchunks * kH * kInPad);
            // parameters, moments and step start at zero
            for (float* q : {net->p, net->m, net->v, net->grad}) {
                hipLaunchKernelGGL(k_mlp_zero_f32, dim3((unsigned)ceil_div(kNParams, 256)), dim3(256), 0, nullptr,
                                   q, kNParams);
                GPE_HIP(hipGetLastError());
            }
            GPE_HIP(hipMemset(net->step, 0, sizeof(int32_t)));
            refresh_copies(net, nullptr);
            GPE_HIP(hipDeviceSynchronize());
        } catch (...) {
            net->release();
            delete net;
            throw;
        }
        *out = net;
    });
}

int gpe_mlp_free(gpe_mlp* net) {
    return mlp_guarded([&] {
        if (!net) return;
        (void)hipDeviceSynchronize();
        net->release();
        delete net;
    });
}

int gpe_mlp_set_params(gpe_mlp* net, const float* params_host, int64_t n) {
    return mlp_guarded([&] {
        require(net != nullptr && params_host != nullptr, "null argument");
        require(n == kNParams, "parameter count mismatch");
        std::vector<float> p(params_host, params_host + n);
        // columns of W1 beyond in_dim must stay zero
        for (int r = 0; r < kH; ++r)
            for (int c = net->in_dim; c < kInPad; ++c)
                require(p[(size_t)(kOffW1 + (int64_t)r * kInPad + c)] == 0.0f,
                        "W1 columns beyond in_dim must be zero");
        GPE_HIP(hipDeviceSynchronize());
        GPE_HIP(hipMemcpy(net->p, p.data(), (size_t)n * sizeof(float), hipMemcpyHostToDevice));
        GPE_HIP(hipMemset(net->m, 0, (size_t)n * sizeof(float)));
        GPE_HIP(hipMemset(net->v, 0, (size_t)n * sizeof(float)));
        GPE_HIP(hipMemset(net->step, 0, sizeof(int32_t)));
        refresh_copies(net, nullptr);
        GPE_HIP(hipDeviceSynchronize());
    });
}

int gpe_mlp_get_params(gpe_mlp* net, float* params_host, int64_t n) {
    return mlp_guarded([&] {
        require(net != nullptr && params_host != nullptr, "null argument");
        require(n == kNParams, "parameter count mismatch");
        GPE_HIP(hipDeviceSynchronize());
        GPE_HIP(hipMemcpy(params_host, net->p, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int gpe_mlp_get_grad(gpe_mlp* net, float* grad_host, int64_t n) {
    return mlp_guarded([&] {
        require(net != nullptr && grad_host != nullptr, "null argument");
        require(n == kNParams, "parameter count mismatch");
        GPE_HIP(hipDeviceSynchronize());
        GPE_HIP(hipMemcpy(grad_host, net->grad, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    });
}

int gpe_mlp_step_count(gpe_mlp* net, int32_t* step_host) {
    return mlp_guarded([&] {
        require(net != nullptr && step_host != nullptr, "null argument");
        GPE_HIP(hipDeviceSynchronize());
        GPE_HIP(hipMemcpy(step_host, net->step, sizeof(int32_t), hipMemcpyDeviceToHost));
    });
}

int gpe_mlp_forward(gpe_mlp* net, const uint16_t* x_dev, int64_t ldx, int64_t rows, float* y_dev, void* stream) {
    return mlp_guarded([&] {
        require(net != nullptr, "null network");
        check_x(net, x_dev, ldx);
        require(rows >= 0, "rows must be >= 0");
        require(y_dev != nullptr || rows == 0, "output is null");
        hipStream_t st = S(stream);
        for (int64_t r0 = 0; r0 < rows; r0 += net->max_rows) {
            const int64_t nr = std::min(net->max_rows, rows - r0);
            forward_chunk(net, x_dev + r0 * ldx, ldx, nr, false, st);
            hipLaunchKernelGGL(k_mlp_head, dim3((unsigned)ceil_div(nr, 256)), dim3(256), 0, st, net->h2, nr,
                               net->w3b, net->p + kOffB3, nullptr, 0.0f, y_dev + r0, nullptr, nullptr);
            GPE_HIP(hipGetLastError());
        }
    });
}

int gpe_mlp_train_step(gpe_mlp* net, const uint16_t* x_dev, int64_t ldx, const uint16_t* xT_dev, int64_t ldxT,
                       const float* target_dev, int64_t rows, float lr, float beta1, float beta2, float eps,
                       float weight_decay, int32_t update, float* loss_dev, void* stream) {
    return mlp_guarded([&] {
        require(net != nullptr, "null network");
        check_x(net, x_dev, ldx);
        require(xT_dev != nullptr && target_dev != nullptr, "null argument");
        require(rows >= 1 && rows <= net->max_rows, "rows must be in [1, max_rows]");
        require(ldxT >= rows && ldxT % 8 == 0, "transposed features need ldxT >= rows, a multiple of 8");
        require((reinterpret_cast<uintptr_t>(xT_dev) & 15) == 0, "transposed features must be 16-byte aligned");
        require(std::isfinite(lr) && lr >= 0.0f && beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f &&
                    eps > 0.0f && weight_decay >= 0.0f,
                "invalid Adam hyperparameters");
        hipStream_t st = S(stream);
        const int64_t tiles = ceil_div(rows, kRowTile);
        const int64_t heads = ceil_div(rows, 256);
        const int64_t wch = ceil_div(rows, kWChunk);
        const float inv_rows = 1.0f / (float)rows;
        forward_chunk(net, x_dev, ldx, rows, true, st);
        hipLaunchKernelGGL(k_mlp_head, dim3((unsigned)heads), dim3(256), 0, st, net->h2, rows, net->w3b,
                           net->p + kOffB3, target_dev, inv_rows, net->y, net->dy, net->loss_part);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_head_bwd, dim3((unsigned)tiles), dim3(256), 0, st, net->dy, net->z2, net->h2, rows,
                           net->w3b, net->dz2, net->dz2T, net->ld_t, net->p_b2, net->p_w3, net->p_b3);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_dh1, dim3((unsigned)tiles), dim3(256), 0, st, net->dz2, rows, net->W2Tb, net->z1,
                           net->dz1T, net->ld_t, net->p_b1);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_wgrad<kH>, dim3((unsigned)wch), dim3(256), 0, st, net->dz2T, net->ld_t, net->h1T,
                           net->ld_t, rows, net->p_W2);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_wgrad<kInPad>, dim3((unsigned)wch), dim3(256), 0, st, net->dz1T, net->ld_t, xT_dev,
                           ldxT, rows, net->p_W1);
        GPE_HIP(hipGetLastError());
        if (loss_dev) {
            hipLaunchKernelGGL(k_mlp_loss, dim3(1), dim3(256), 0, st, net->loss_part, heads, inv_rows, loss_dev);
            GPE_HIP(hipGetLastError());
        }
        AdamArgs a{};
        a.pc[0] = Piece{net->p_W1, wch, kOffB1 - kOffW1, kOffW1};
        a.pc[1] = Piece{net->p_b1, tiles, kH, kOffB1};
        a.pc[2] = Piece{net->p_W2, wch, kOffB2 - kOffW2, kOffW2};
        a.pc[3] = Piece{net->p_b2, tiles, kH, kOffB2};
        a.pc[4] = Piece{net->p_w3, tiles, kH, kOffW3};
        a.pc[5] = Piece{net->p_b3, tiles, 1, kOffB3};
        a.p = net->p;
        a.m = net->m;
        a.v = net->v;
        a.grad = net->grad;
        a.W1b = net->W1b;
        a.W2b = net->W2b;
        a.W2Tb = net->W2Tb;
        a.w3b = net->w3b;
        a.step = net->step;
        a.lr = lr;
        a.beta1 = beta1;
        a.beta2 = beta2;
        a.eps = eps;
        a.wd = weight_decay;
        a.update = update ? 1 : 0;
        hipLaunchKernelGGL(k_mlp_adam, dim3((unsigned)ceil_div(kNParams, 256)), dim3(256), 0, st, a);
        GPE_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_mlp_step_inc, dim3(1), dim3(64), 0, st, net->step, update ? 1 : 0);
        GPE_HIP(hipGetLastError());
    });
}

}  // extern "C"
