// Per-SNP feature algebra of the GxE per-block backend (DirectContext,
// src/native/gxeldcore.cpp:8033-8243): from the weighted sums of squares and
// the design-basis moments of each standardized genotype column, the projected
// feature scales, norms, environment-weighted diagonals, the x-w correlation
// and the projection-leakage ratios, in FP64, one thread per SNP.  The inputs
// come from the engine's own passes (gpe_weighted_sumsq, gpe_xtv_f64), so no
// projected N x L panel is ever formed.
#include "../../include/summit_gpe.h"
#include "gpe_internal.h"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstring>
#include <string>

using namespace gpe;

namespace {

// s: 4 x L rows (weights 1, e, e^2, e^4) of the engine's x^2 sums; u: L x 4q moments
// [Q, EQ, E^2Q, E^3Q]^T x already in the DirectContext standardization; r2: (L,) the
// squared per-SNP factor between the two standardizations; gram, e2gram: q x q
// (gram[b q + a] = Q_a . Q_b, e2gram[b q + a] = Q_a . (e^2 Q_b)).
__global__ __launch_bounds__(256) void k_gxe_direct_features(const double* __restrict__ s, int64_t lds,
                                                             const double* __restrict__ u, int64_t ldu, int q,
                                                             const double* __restrict__ gram,
                                                             const double* __restrict__ e2gram,
                                                             const double* __restrict__ r2, double rank, int64_t L,
                                                             double* __restrict__ out, int64_t ldo) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= L) return;
    const double f = r2[j];
    const double s0 = s[j] * f, s1 = s[lds + j] * f, s2 = s[2 * lds + j] * f, s4 = s[3 * lds + j] * f;
    const double* uj = u + j * ldu;  // the SNP's 4q moments, read from L1/L2 (no private arrays)
    double u00 = 0.0, u11 = 0.0, u01 = 0.0, u0u2 = 0.0, u1u3 = 0.0;
    for (int a = 0; a < q; ++a) {
        const double v0 = uj[a], v1 = uj[q + a], v2 = uj[2 * q + a], v3 = uj[3 * q + a];
        u00 += v0 * v0;
        u11 += v1 * v1;
        u01 += v0 * v1;
        u0u2 += v0 * v2;
        u1u3 += v1 * v3;
    }
    double u0e2u0 = 0.0, u1e2u1 = 0.0, leak_x = 0.0, leak_w = 0.0;
    for (int a = 0; a < q; ++a) {
        double e2u0 = 0.0, e2u1 = 0.0, g0 = 0.0, g1 = 0.0;
        for (int b = 0; b < q; ++b) {
            const double e2 = e2gram[(int64_t)b * q + a], g = gram[(int64_t)b * q + a];
            const double v0 = uj[b], v1 = uj[q + b];
            e2u0 += e2 * v0;
            e2u1 += e2 * v1;
            g0 += g * v0;
            g1 += g * v1;
        }
        const double v0 = uj[a], v1 = uj[q + a];
        u0e2u0 += v0 * e2u0;
        u1e2u1 += v1 * e2u1;
        const double lx = v0 - g0, lw = v1 - g1;
        leak_x += lx * lx;
        leak_w += lw * lw;
    }
    const double ssx = s0 - u00, ssw = s2 - u11;
    const double varx = ssx / rank, varw = ssw / rank;
    const double sx = 1.0 / sqrt(varx), sw = 1.0 / sqrt(varw);
    out[j] = sx;
    out[ldo + j] = sw;
    out[2 * ldo + j] = sx * sx * ssx / rank;
    out[3 * ldo + j] = sw * sw * ssw / rank;
    out[4 * ldo + j] = sx * sx * (s2 - 2.0 * u0u2 + u0e2u0) / rank;
    out[5 * ldo + j] = sw * sw * (s4 - 2.0 * u1u3 + u1e2u1) / rank;
    out[6 * ldo + j] = sx * sw * (s1 - u01) / rank;
    out[7 * ldo + j] = varx;
    out[8 * ldo + j] = varw;
    out[9 * ldo + j] = sqrt(fmax(leak_x, 0.0) / fmax(ssx, DBL_MIN));
    out[10 * ldo + j] = sqrt(fmax(leak_w, 0.0) / fmax(ssw, DBL_MIN));
}

template <class F>
int gxe_guarded(F&& fn) {
    try {
        fn();
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

void gxe_require(bool ok, const char* msg) {
    if (!ok) throw Error(msg);
}

}  // namespace

extern "C" {

int gpe_gxe_direct_features(const double* sums_dev, int64_t lds, const double* moments_dev, int64_t ldm,
                            int32_t q, const double* gram_dev, const double* e2gram_dev, const double* r2_dev,
                            double rank, int64_t n_snps, double* out_dev, int64_t ldo, void* stream) {
    return gxe_guarded([&] {
        gxe_require(q >= 1, "design basis rank must be positive");
        gxe_require(n_snps >= 0 && lds >= n_snps && ldm >= 4 * (int64_t)q && ldo >= n_snps,
                    "bad feature-algebra leading dimension");
        gxe_require(rank > 0.0, "residual rank must be positive");
        if (n_snps == 0) return;
        gxe_require(sums_dev && moments_dev && gram_dev && e2gram_dev && r2_dev && out_dev, "null argument");
        hipStream_t st = reinterpret_cast<hipStream_t>(stream);
        hipLaunchKernelGGL(k_gxe_direct_features, dim3((unsigned)((n_snps + 255) / 256)), dim3(256), 0, st, sums_dev,
                           lds, moments_dev, ldm, q, gram_dev, e2gram_dev, r2_dev, rank, n_snps, out_dev, ldo);
        GPE_HIP(hipGetLastError());
    });
}

}  // extern "C"
