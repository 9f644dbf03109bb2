// C ABI of libsummit_gpe.so (declared in include/summit_gpe.h).
//
// Host orchestration of the genotype-probe engine: device residency of the
// genotype matrix, per-SNP statistics, probe streams, and the two LD-score
// phases built from the packed-contraction kernel in gpe_kernels.hip.
#include "../../include/summit_gpe.h"
#include "gpe_internal.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <random>
#include <string>
#include <thread>
#include <vector>

using namespace gpe;

namespace {

thread_local std::string g_err;

struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
};

}  // namespace

struct gpe_geno {
    int device = 0;
    int64_t n = 0, m = 0;
    int64_t snp_stride = 0, smp_stride = 0;
    int64_t m_alloc = 0, n_alloc = 0;
    uint8_t* snp_raw = nullptr;
    uint8_t* snp_work = nullptr;  // == snp_raw unless HWE-completed
    uint8_t* smp_work = nullptr;
    // chunk-tiled copies of snp_work / smp_work read by the contractions (built by prepare)
    uint8_t* snp_tiled = nullptr;
    uint8_t* smp_tiled = nullptr;
    bool prepared = false;
    int ddof = 1;
    int impute_mode = GPE_IMPUTE_MEAN;
    uint64_t impute_seed = 0;
    int32_t* nobs = nullptr;
    int64_t* sum = nullptr;
    int64_t* sumsq = nullptr;
    double* mean = nullptr;
    double* inv_sd = nullptr;
    int64_t n_missing = 0;
    int64_t* miss_ptr = nullptr;   // per sample: sorted SNP indices of its missing calls
    int32_t* miss_idx = nullptr;
    int64_t* smiss_ptr = nullptr;  // per SNP: sorted sample indices of its missing calls
    int32_t* smiss_idx = nullptr;
    int64_t* smiss_rptr = nullptr; // per SNP: list offsets of rptr_parts sample ranges
    int rptr_parts = 0;
    std::map<std::string, Buf> ws;
    // set while an entry point enqueues onto a stream under HIP graph capture:
    // workspaces may then not grow (no allocation may enter a graph)
    bool capturing = false;
    // device copy of the phase-1 SNP-block starts for (begin, end, step)
    int64_t starts_key[3] = {-1, -1, -1};
    // pinned host staging of host-generated (real-valued) probes, reused
    double* host_stage = nullptr;
    size_t host_stage_bytes = 0;

    double* stage(size_t bytes) {
        if (host_stage_bytes < bytes) {
            if (host_stage) GPE_HIP(hipHostFree(host_stage));
            host_stage = nullptr;
            host_stage_bytes = 0;
            GPE_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_stage), bytes, hipHostMallocDefault));
            host_stage_bytes = bytes;
        }
        return host_stage;
    }

    void* scratch(const std::string& name, size_t bytes) {
        Buf& b = ws[name];
        if (b.bytes < bytes) {
            if (capturing)
                throw Error("workspace '" + name +
                            "' is not sized for this call: run the same call once outside graph capture first");
            if (b.p) GPE_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
            GPE_HIP(hipMalloc(&b.p, std::max<size_t>(bytes, 256)));
            b.bytes = std::max<size_t>(bytes, 256);
        }
        return b.p;
    }
};

namespace {

template <class F>
int guarded(F&& f) {
    try {
        f();
        return GPE_OK;
    } catch (const Error& e) {
        g_err = e.what();
        return std::strncmp(e.what(), "HIP error", 9) == 0 ? GPE_EDEVICE : GPE_EINVAL;
    } catch (const std::exception& e) {
        g_err = e.what();
        return GPE_EINTERNAL;
    } catch (...) {
        g_err = "unknown error";
        return GPE_EINTERNAL;
    }
}

inline hipStream_t S(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Marks the handle as capturing while an entry point enqueues onto a stream
// under HIP graph capture.
struct CaptureScope {
    gpe_geno* g;
    bool active = false;
    CaptureScope(gpe_geno* g_, hipStream_t st) : g(g_) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GPE_HIP(hipStreamIsCapturing(st, &cs));
        active = cs == hipStreamCaptureStatusActive;
        if (g) g->capturing = active;
    }
    ~CaptureScope() {
        if (g) g->capturing = false;
    }
};

}  // namespace

void gpe::set_last_error(const char* msg) { g_err = msg ? msg : ""; }

namespace {

void require(bool ok, const std::string& msg) {
    if (!ok) throw Error(msg);
}

int bitlen(uint64_t x) {
    int b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b;
}

template <class T>
T* dalloc(int64_t count) {
    void* p = nullptr;
    GPE_HIP(hipMalloc(&p, std::max<int64_t>(count, 1) * sizeof(T)));
    return static_cast<T*>(p);
}

void dfree(void* p) {
    if (p) (void)hipFree(p);
}

// w[i][k] = inv_sd_i * inv_i * sqrt(a_ik), 0 where phase 1 skips the entry
// (src/native/gwldcore.cpp:1002-1063: a == 0, a <= 0 or non-finite, inv <= 0
// or non-finite, non-finite scale).
// True when every standardized value of SNP j is 0: no calls, or every call
// equals the mean (cnt * sumsq == sum^2, exact in integers).  Such a column
// contributes nothing to any product, so its weight is set to 0; a huge weight
// (e.g. the 1/sqrt(eps) residual scale of a monomorphic SNP) would otherwise
// set the fixed-point scale of its whole coefficient column.
__device__ __forceinline__ bool zero_column(const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t j) {
    const __int128 cnt = nobs[j];
    const __int128 sm = sum[j];
    return cnt == 0 || cnt * (__int128)sumsq[j] == sm * sm;
}

__global__ void k_make_w(const double* annot, int n_bins, const double* inv, const double* inv_sd,
                         const int32_t* nobs, const int64_t* sum, const int64_t* sumsq, int64_t k0, int64_t k1,
                         double* w) {
    const int64_t i = k0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k1) return;
    const double iv = zero_column(nobs, sum, sumsq, i) ? 0.0 : inv[i - k0];
    for (int k = 0; k < n_bins; ++k) {
        const double a = annot[(i - k0) * n_bins + k];
        double out = 0.0;
        if (a != 0.0 && a > 0.0 && isfinite(a) && isfinite(iv) && iv > 0.0) {
            const double sc = iv * sqrt(a);
            if (isfinite(sc)) out = inv_sd[i] * sc;
        }
        w[i * n_bins + k] = out;
    }
}

__global__ void k_absmax_i32(const int32_t* u, int64_t ldu, int64_t n, int n_cols, unsigned int* out) {
    const int c = blockIdx.y;
    unsigned int mx = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = u[(int64_t)c * ldu + i];
        const unsigned int a = v < 0 ? (unsigned int)(-(int64_t)v) : (unsigned int)v;
        mx = mx > a ? mx : a;
    }
    atomicMax(out, mx);
}

void free_prepared(gpe_geno* g) {
    if (g->snp_work && g->snp_work != g->snp_raw) dfree(g->snp_work);
    g->snp_work = g->snp_raw;
    dfree(g->snp_tiled);
    dfree(g->smp_tiled);
    g->snp_tiled = nullptr;
    g->smp_tiled = nullptr;
    dfree(g->miss_ptr);
    dfree(g->miss_idx);
    dfree(g->smiss_ptr);
    dfree(g->smiss_idx);
    dfree(g->smiss_rptr);
    g->smiss_rptr = nullptr;
    g->rptr_parts = 0;
    g->miss_ptr = nullptr;
    g->miss_idx = nullptr;
    g->smiss_ptr = nullptr;
    g->smiss_idx = nullptr;
    g->n_missing = 0;
    g->prepared = false;
}

void require_prepared(const gpe_geno* g) {
    require(g != nullptr, "null genotype handle");
    require(g->prepared, "genotype handle is not prepared (call gpe_geno_prepare)");
}

struct LimbCfg {
    int nl, ctc;
};
LimbCfg limb_cfg(int limbs) {
    if (limbs == 5) return {5, 1};
    if (limbs == 6) return {6, 1};
    throw Error("limbs must be 5 or 6");
}

// Optional HIP-event timing of the engine's kernels, recorded on the stream
// each kernel is launched on (gpe_profile_*).
struct ProfEvent {
    int kind;
    hipEvent_t a, b;
};
struct Profiler {
    bool on = false;
    std::vector<ProfEvent> pending;
    double ms[GPE_PROF_KINDS] = {};
    int64_t count[GPE_PROF_KINDS] = {};
};
Profiler& prof() {
    static Profiler p;
    return p;
}
struct ProfScope {
    int kind;
    hipStream_t st;
    hipEvent_t a = nullptr, b = nullptr;
    ProfScope(int k, hipStream_t s) : kind(k), st(s) {
        if (!prof().on) return;
        GPE_HIP(hipEventCreate(&a));
        GPE_HIP(hipEventCreate(&b));
        GPE_HIP(hipEventRecord(a, st));
    }
    void stop() {
        if (!a) return;
        GPE_HIP(hipEventRecord(b, st));
        prof().pending.push_back({kind, a, b});
        a = nullptr;
    }
};

// K splits of the MFMA-bound phase-1 contraction: workgroups of equal work
// finish unevenly, so the last wave's idle slots cost less the more waves
// there are.  Add splits until the grid holds 16 waves of resident
// workgroups, while every split keeps >= 128 chunks (16,384 SNPs) so that
// prologue and epilogue stay small, at most 16.  Measured at N=100,000,
// M=500,000, B=64 (782 workgroups per split, 512 slots): 11.5 ms at 11-16
// splits against 12.3 ms at the 3 that pick_splits' fill rule gives; one split
// at N=5,000, M=10,000 (79 chunks).  Results do not depend on the count.
int pick_splits_deep(int64_t grid_xy, int64_t kcount, int64_t max_chunks_per_split, int64_t slots) {
    int64_t s = 1;
    while (s < 16 && grid_xy * s < 16 * slots && kcount / (s + 1) >= 128) ++s;
    while (ceil_div(kcount, s) > max_chunks_per_split) ++s;
    return (int)s;
}

// Real-valued probes of blocks [starts[b], next start or blk_end) for V-tile
// v_start, written to dst[(row - starts[0]) * ldz + c] for c < v_count; every
// block's stream is independent, so blocks are spread over host threads.
void real_probes_into(int rand_dist, uint64_t root, const int64_t* starts, int n_blocks, int64_t blk_end, int v_start,
                      int v_count, double* dst, int64_t ldz);

// Packed rows [src, src + rows * stride) to host rows of host_stride bytes (the
// first `used` bytes of each): whole rows in 64 MiB pieces through the pinned
// staging buffer, then compacted on the host (plain contiguous copies only).
void download_rows_staged(gpe_geno* g, uint8_t* host, int64_t host_stride, const uint8_t* src, int64_t stride,
                          int64_t used, int64_t rows) {
    GPE_HIP(hipDeviceSynchronize());
    const int64_t piece = std::max<int64_t>(1, ((int64_t)64 << 20) / stride);
    uint8_t* stage = reinterpret_cast<uint8_t*>(g->stage((size_t)(std::min(rows, piece) * stride)));
    for (int64_t r0 = 0; r0 < rows; r0 += piece) {
        const int64_t nr = std::min(piece, rows - r0);
        GPE_HIP(hipMemcpy(stage, src + r0 * stride, (size_t)(nr * stride), hipMemcpyDeviceToHost));
        for (int64_t r = 0; r < nr; ++r) std::memcpy(host + (r0 + r) * host_stride, stage + r * stride, (size_t)used);
    }
}

bool is_real_probe_dist(int d) {
    return d == GPE_RAND_GAUSSIAN || d == GPE_RAND_SPHERICAL || d == GPE_RAND_GAUSSIAN_F32 ||
           d == GPE_RAND_SPHERICAL_F32;
}

// The reference's real-valued probe fill for one SNP block
// (src/native/gwldcore.cpp:1679-1716): std::mt19937_64 seeded by
// make_seed(root, block start, v_start), one std::normal_distribution<T> (PolarNormal) for the
// whole block, column by column; spherical probes rescale each column to norm
// sqrt(L) with the sum of squares in long double.  T is the precision of the
// caller's arrays (float32 callers draw float variates).  Writes z[r * ldz + c].
// std::normal_distribution<T> over std::mt19937_64 as libstdc++ defines it
// (Marsaglia's polar method; each uniform is generate_canonical<T, digits>, one
// 64-bit draw u giving T(u) / 2^64, clamped below 1; the second variate of an
// accepted pair is kept for the next call), written out so that the constant
// log2 of the engine range, which libstdc++ evaluates in long double at every
// uniform unless the compiler folds it, is not recomputed per draw.
template <class T>
struct PolarNormal {
    bool saved_available = false;
    T saved = T(0);
    static T canonical(std::mt19937_64& rng) {
        const T ret = T(rng() - std::mt19937_64::min()) / T(18446744073709551616.0L);
        return ret >= T(1) ? std::nextafter(T(1), T(0)) : ret;
    }
    T operator()(std::mt19937_64& rng) {
        T ret;
        if (saved_available) {
            saved_available = false;
            ret = saved;
        } else {
            T x, y, r2;
            do {
                x = T(2.0) * canonical(rng) - 1.0;
                y = T(2.0) * canonical(rng) - 1.0;
                r2 = x * x + y * y;
            } while (r2 > 1.0 || r2 == 0.0);
            const T mult = std::sqrt(-2 * std::log(r2) / r2);
            saved = x * mult;
            saved_available = true;
            ret = y * mult;
        }
        return ret * T(1) + T(0);
    }
};

template <class T>
void real_probe_fill(uint64_t root, int64_t blk_start, int v_start, int64_t L, int V, bool spherical, double* z,
                     int64_t ldz) {
    auto mix = [](uint64_t x) {
        x += 0x9e3779b97f4a7c15ULL;
        x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
        x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
        return x ^ (x >> 31);
    };
    uint64_t seed = 0x1234abcdULL;
    seed ^= mix(root);
    seed ^= mix((uint64_t)(int64_t)(int32_t)blk_start + 0x9e37ULL);
    seed ^= mix((uint64_t)(int64_t)v_start + 0x85ebULL);
    std::mt19937_64 rng(seed);
    PolarNormal<T> gN;
    std::vector<T> col((size_t)L);
    for (int c = 0; c < V; ++c) {
        long double ss = 0.0L;
        for (int64_t r = 0; r < L; ++r) {
            const T v = gN(rng);
            col[(size_t)r] = v;
            if (spherical) ss += (long double)v * (long double)v;
        }
        if (spherical) {
            const T sc = ss > 0.0L ? (T)std::sqrt((long double)L / ss) : T(1);
            for (int64_t r = 0; r < L; ++r) col[(size_t)r] *= sc;
        }
        for (int64_t r = 0; r < L; ++r) z[r * ldz + c] = (double)col[(size_t)r];
    }
}

// memcpy split over n_threads host threads (page-cache or mmap reads of a BED
// body are far slower from one thread than the PCIe copy that follows)
void parallel_memcpy(uint8_t* dst, const uint8_t* src, int64_t bytes, int n_threads) {
    if (n_threads <= 1 || bytes < ((int64_t)4 << 20)) {
        std::memcpy(dst, src, (size_t)bytes);
        return;
    }
    std::vector<std::thread> pool;
    const int64_t part = ceil_div(bytes, n_threads);
    for (int t = 0; t < n_threads; ++t) {
        const int64_t b0 = t * part, b1 = std::min(bytes, b0 + part);
        if (b0 >= b1) break;
        pool.emplace_back([=] { std::memcpy(dst + b0, src + b0, (size_t)(b1 - b0)); });
    }
    for (auto& th : pool) th.join();
}

// Schedule overrides written by the autotuner: (knob, rows, length, columns,
// limbs) -> value.  Read on every traversal; guarded by a mutex.
struct SchedKey {
    int32_t knob;
    int64_t rows, length;
    int32_t columns, limbs;
    bool operator<(const SchedKey& o) const {
        return std::tie(knob, rows, length, columns, limbs) < std::tie(o.knob, o.rows, o.length, o.columns, o.limbs);
    }
};
std::mutex& sched_mu() {
    static std::mutex m;
    return m;
}
std::map<SchedKey, int32_t>& sched_table() {
    static std::map<SchedKey, int32_t> t;
    return t;
}
int32_t sched_lookup(int32_t knob, int64_t rows, int64_t length, int32_t columns, int32_t limbs) {
    std::lock_guard<std::mutex> lock(sched_mu());
    auto& t = sched_table();
    auto it = t.find(SchedKey{knob, rows, length, columns, limbs});
    return it == t.end() ? 0 : it->second;
}

// K splits of a contraction whose partial sums are added atomically.  Equal-work
// workgroups finish in rounds of `slots` (the device's resident capacity), so the
// split count is the smallest one whose last round is at least 90% full (the best
// such fill up to 16 splits otherwise), then raised until no split exceeds
// max_chunks_per_split.
int pick_splits(int64_t grid_xy, int64_t kcount, int64_t max_chunks_per_split, int64_t slots) {
    int64_t best = 1;
    double best_fill = 0.0;
    for (int64_t s = 1; s <= 16 && s <= kcount; ++s) {
        const int64_t blocks = grid_xy * s;
        const double fill = (double)blocks / (double)(ceil_div(blocks, slots) * slots);
        if (fill > best_fill + 1e-9) {
            best_fill = fill;
            best = s;
        }
        if (fill >= 0.9) break;
    }
    while (ceil_div(kcount, best) > max_chunks_per_split) ++best;
    return (int)best;
}

void real_probes_into(int rand_dist, uint64_t root, const int64_t* starts, int n_blocks, int64_t blk_end, int v_start,
                      int v_count, double* dst, int64_t ldz) {
    const int64_t s0 = starts[0];
    auto fill = [&](int b) {
        const int64_t s = starts[b];
        const int64_t e = b + 1 < n_blocks ? starts[b + 1] : blk_end;
        if (rand_dist >= GPE_RAND_GAUSSIAN_F32)
            real_probe_fill<float>(root, s, v_start, e - s, v_count, rand_dist == GPE_RAND_SPHERICAL_F32,
                                   dst + (s - s0) * ldz, ldz);
        else
            real_probe_fill<double>(root, s, v_start, e - s, v_count, rand_dist == GPE_RAND_SPHERICAL,
                                    dst + (s - s0) * ldz, ldz);
    };
    const int hw = (int)std::max(1u, std::thread::hardware_concurrency());
    const int nt = std::min({hw, 16, n_blocks});
    if (nt <= 1 || (blk_end - s0) * (int64_t)v_count < (1 << 16)) {
        for (int b = 0; b < n_blocks; ++b) fill(b);
        return;
    }
    std::vector<std::thread> pool;
    for (int t = 0; t < nt; ++t)
        pool.emplace_back([&, t] {
            for (int b = t; b < n_blocks; b += nt) fill(b);
        });
    for (auto& th : pool) th.join();
}

}  // namespace

extern "C" {

const char* gpe_last_error(void) { return g_err.c_str(); }

int gpe_profile_enable(int32_t on) {
    return guarded([&] { prof().on = on != 0; });
}

int gpe_profile_read(int32_t kind, double* total_ms, int64_t* launches) {
    return guarded([&] {
        require(kind >= 0 && kind < GPE_PROF_KINDS, "unknown profile kind");
        Profiler& p = prof();
        for (auto& e : p.pending) {
            GPE_HIP(hipEventSynchronize(e.b));
            float ms = 0.f;
            GPE_HIP(hipEventElapsedTime(&ms, e.a, e.b));
            p.ms[e.kind] += ms;
            p.count[e.kind] += 1;
            (void)hipEventDestroy(e.a);
            (void)hipEventDestroy(e.b);
        }
        p.pending.clear();
        if (total_ms) *total_ms = p.ms[kind];
        if (launches) *launches = p.count[kind];
    });
}

int gpe_schedule_set(int32_t knob, int64_t rows, int64_t length, int32_t columns, int32_t limbs, int32_t value) {
    return guarded([&] {
        require(knob >= 0 && knob < GPE_SCHED_KNOBS, "unknown schedule knob");
        require(rows > 0 && length > 0 && columns > 0 && limbs > 0, "schedule keys must be positive");
        if (knob == GPE_SCHED_P1_SPLITS) require(value <= 64, "phase-1 splits must be in [1, 64]");
        if (knob == GPE_SCHED_P1_FINALIZE_KIB)
            require(value <= 0 || (value >= 64 && value <= 65536), "finalize range must be in [64, 65536] KiB");
        if (knob == GPE_SCHED_P1_CPB || knob == GPE_SCHED_P2_CPB)
            require(value <= 2, "chunks per barrier must be 1 or 2");
        if (knob == GPE_SCHED_P2_MISS || knob == GPE_SCHED_P1_MISS)
            require(value <= 2, "missing-call method must be 1 or 2");
        if (knob == GPE_SCHED_P2_MISS_PARTS)
            require(value <= 0 || value == 1 || (value % 8 == 0 && value <= 64),
                    "missing-call sample ranges must be 1 or a multiple of 8 up to 64");
        std::lock_guard<std::mutex> lock(sched_mu());
        const SchedKey k{knob, rows, length, columns, limbs};
        if (value <= 0)
            sched_table().erase(k);
        else
            sched_table()[k] = value;
    });
}

int gpe_schedule_get(int32_t knob, int64_t rows, int64_t length, int32_t columns, int32_t limbs, int32_t* value) {
    return guarded([&] {
        require(knob >= 0 && knob < GPE_SCHED_KNOBS, "unknown schedule knob");
        require(value != nullptr, "value is null");
        *value = sched_lookup(knob, rows, length, columns, limbs);
    });
}

int gpe_schedule_clear(void) {
    return guarded([&] {
        std::lock_guard<std::mutex> lock(sched_mu());
        sched_table().clear();
    });
}

int gpe_schedule_split_rule(int32_t rule, int64_t grid_xy, int64_t kcount, int64_t max_chunks_per_split,
                            int64_t slots, int32_t* splits) {
    return guarded([&] {
        require(splits != nullptr && grid_xy > 0 && kcount > 0 && max_chunks_per_split > 0 && slots > 0,
                "split rule needs positive sizes");
        require(rule == 0 || rule == 1, "split rule must be 0 (fill) or 1 (deep)");
        *splits = rule == 0 ? pick_splits(grid_xy, kcount, max_chunks_per_split, slots)
                            : pick_splits_deep(grid_xy, kcount, max_chunks_per_split, slots);
    });
}

int gpe_profile_reset(void) {
    return guarded([&] {
        double ms;
        int64_t n;
        if (gpe_profile_read(0, &ms, &n) != 0) throw Error(g_err);
        Profiler& p = prof();
        for (int k = 0; k < GPE_PROF_KINDS; ++k) {
            p.ms[k] = 0.0;
            p.count[k] = 0;
        }
    });
}

const char* gpe_version(void) { return "summit-gpe 0.1 (gfx950)"; }

int gpe_set_device(int device) {
    return guarded([&] { GPE_HIP(hipSetDevice(device)); });
}

int gpe_geno_create(int64_t n_samples, int64_t n_snps, gpe_geno** out) {
    return guarded([&] {
        require(out != nullptr, "null output handle");
        require(n_samples > 1 && n_snps > 0, "genotype matrix needs n_samples > 1 and n_snps > 0");
        require(n_samples < (int64_t)1 << 31 && n_snps < (int64_t)1 << 31, "genotype matrix too large");
        auto* g = new gpe_geno();
        try {
            GPE_HIP(hipGetDevice(&g->device));
            g->n = n_samples;
            g->m = n_snps;
            g->snp_stride = round_up(ceil_div(n_samples, 4), 64);
            g->smp_stride = round_up(ceil_div(n_snps, 4), 64);
            g->m_alloc = round_up(n_snps, kRowPad) + kRowPad;
            g->n_alloc = round_up(n_samples, kRowPad) + kRowPad;
            g->snp_raw = dalloc<uint8_t>(g->m_alloc * g->snp_stride);
            g->snp_work = g->snp_raw;
            g->smp_work = dalloc<uint8_t>(g->n_alloc * g->smp_stride);
            launch_fill_u8(g->snp_raw, g->m_alloc * g->snp_stride, 0, nullptr);
            launch_fill_u8(g->smp_work, g->n_alloc * g->smp_stride, 0, nullptr);
            g->nobs = dalloc<int32_t>(g->m_alloc);
            g->sum = dalloc<int64_t>(g->m_alloc);
            g->sumsq = dalloc<int64_t>(g->m_alloc);
            g->mean = dalloc<double>(g->m_alloc);
            g->inv_sd = dalloc<double>(g->m_alloc);
            launch_fill_u8(reinterpret_cast<uint8_t*>(g->mean), g->m_alloc * 8, 0, nullptr);
            launch_fill_u8(reinterpret_cast<uint8_t*>(g->inv_sd), g->m_alloc * 8, 0, nullptr);
            GPE_HIP(hipDeviceSynchronize());
        } catch (...) {
            gpe_geno_free(g);
            throw;
        }
        *out = g;
    });
}

int gpe_geno_free(gpe_geno* g) {
    return guarded([&] {
        if (!g) return;
        free_prepared(g);
        dfree(g->snp_raw);
        dfree(g->smp_work);
        dfree(g->nobs);
        dfree(g->sum);
        dfree(g->sumsq);
        dfree(g->mean);
        dfree(g->inv_sd);
        for (auto& kv : g->ws) dfree(kv.second.p);
        if (g->host_stage) (void)hipHostFree(g->host_stage);
        delete g;
    });
}

int gpe_geno_shape(gpe_geno* g, int64_t* n_samples, int64_t* n_snps, int64_t* n_missing) {
    return guarded([&] {
        require(g != nullptr, "null genotype handle");
        if (n_samples) *n_samples = g->n;
        if (n_snps) *n_snps = g->m;
        if (n_missing) *n_missing = g->n_missing;
    });
}

int gpe_geno_upload_bed(gpe_geno* g, const uint8_t* bed_host, int64_t n_total, int64_t host_stride,
                        int64_t snp_begin, int64_t snp_count, const int64_t* rows_host, void* stream) {
    return guarded([&] {
        require(g != nullptr && bed_host != nullptr, "null argument to gpe_geno_upload_bed");
        require(host_stride >= ceil_div(n_total, 4), "host_stride smaller than ceil(n_total/4)");
        require(snp_begin >= 0 && snp_count >= 0 && snp_begin + snp_count <= g->m, "SNP range out of bounds");
        if (!rows_host) require(n_total == g->n, "without rows the BED sample count must equal n_samples");
        hipStream_t st = S(stream);
        int64_t* rows_dev = nullptr;
        if (rows_host) {
            for (int64_t i = 0; i < g->n; ++i)
                require(rows_host[i] >= 0 && rows_host[i] < n_total,
                        "row_sel contains an out-of-range sample index");
            rows_dev = static_cast<int64_t*>(g->scratch("upload_rows", g->n * sizeof(int64_t)));
            GPE_HIP(hipMemcpyAsync(rows_dev, rows_host, g->n * sizeof(int64_t), hipMemcpyHostToDevice, st));
        }
        // Stream through two pinned staging buffers in chunks of SNP rows: while the
        // device copies and scatters chunk j, host threads fill the other buffer
        // with chunk j+1.  An event per buffer marks when its last device use ended.
        const int64_t chunk_rows = std::max<int64_t>(1, (int64_t)(64 << 20) / std::max<int64_t>(host_stride, 1));
        const int64_t chunk_bytes = chunk_rows * host_stride;
        uint8_t* pinned[2] = {nullptr, nullptr};
        hipEvent_t done[2] = {nullptr, nullptr};
        auto release = [&] {
            for (int b = 0; b < 2; ++b) {
                if (done[b]) (void)hipEventDestroy(done[b]);
                if (pinned[b]) (void)hipHostFree(pinned[b]);
            }
        };
        try {
            for (int b = 0; b < 2; ++b) {
                GPE_HIP(hipHostMalloc(reinterpret_cast<void**>(&pinned[b]), chunk_bytes, hipHostMallocDefault));
                GPE_HIP(hipEventCreateWithFlags(&done[b], hipEventDisableTiming));
            }
            uint8_t* staging[2] = {static_cast<uint8_t*>(g->scratch("upload_staging0", chunk_bytes)),
                                   static_cast<uint8_t*>(g->scratch("upload_staging1", chunk_bytes))};
            const int n_threads = (int)std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
            int64_t j = 0;
            for (int64_t j0 = 0; j0 < snp_count; j0 += chunk_rows, ++j) {
                const int b = (int)(j & 1);
                const int64_t rows_now = std::min(chunk_rows, snp_count - j0);
                const int64_t bytes = rows_now * host_stride;
                if (j >= 2) GPE_HIP(hipEventSynchronize(done[b]));
                parallel_memcpy(pinned[b], bed_host + j0 * host_stride, bytes, n_threads);
                GPE_HIP(hipMemcpyAsync(staging[b], pinned[b], bytes, hipMemcpyHostToDevice, st));
                uint8_t* dst = g->snp_raw + (snp_begin + j0) * g->snp_stride;
                if (rows_dev) {
                    launch_select_rows(staging[b], host_stride, rows_now, rows_dev, g->n, dst, g->snp_stride, st);
                } else {
                    GPE_HIP(hipMemcpy2DAsync(dst, g->snp_stride, staging[b], host_stride, ceil_div(n_total, 4),
                                             rows_now, hipMemcpyDeviceToDevice, st));
                }
                launch_clear_tail(dst, rows_now, g->snp_stride, g->n, st);
                GPE_HIP(hipEventRecord(done[b], st));
            }
            GPE_HIP(hipStreamSynchronize(st));
        } catch (...) {
            (void)hipStreamSynchronize(st);
            release();
            throw;
        }
        release();
        free_prepared(g);
    });
}

int gpe_geno_synthesize(gpe_geno* g, uint64_t seed, double maf_lo, double maf_hi, double miss_rate,
                        int32_t block_len, double copy_prob, void* stream) {
    return guarded([&] {
        require(g != nullptr, "null genotype handle");
        require(block_len > 0, "block_len must be positive");
        require(maf_lo >= 0.0 && maf_hi <= 1.0 && maf_lo <= maf_hi, "MAF range must lie in [0, 1]");
        require(miss_rate >= 0.0 && miss_rate < 1.0, "miss_rate must lie in [0, 1)");
        launch_synth_markov(g->snp_raw, g->snp_stride, g->m, g->n, seed, maf_lo, maf_hi, miss_rate, block_len,
                            copy_prob, copy_prob <= 0.0 ? 1 : 0, S(stream));
        free_prepared(g);
    });
}

int gpe_geno_prepare(gpe_geno* g, int32_t ddof, int32_t impute_mode, uint64_t impute_seed, int64_t snp_offset,
                     const int64_t* global_rows_host, void* stream) {
    return guarded([&] {
        require(g != nullptr, "null genotype handle");
        require(impute_mode == GPE_IMPUTE_MEAN || impute_mode == GPE_IMPUTE_HWE,
                "impute_mode must be 'mean' or 'hwe'");
        hipStream_t st = S(stream);
        free_prepared(g);
        g->ddof = ddof;
        g->impute_mode = impute_mode;
        g->impute_seed = impute_seed;
        launch_snp_stats(g->snp_raw, g->snp_stride, g->m, g->n, g->nobs, g->sum, g->sumsq, st);
        if (impute_mode == GPE_IMPUTE_HWE) {
            g->snp_work = dalloc<uint8_t>(g->m_alloc * g->snp_stride);
            GPE_HIP(hipMemcpyAsync(g->snp_work, g->snp_raw, g->m_alloc * g->snp_stride, hipMemcpyDeviceToDevice, st));
            int64_t* rows_dev = nullptr;
            if (global_rows_host) {
                rows_dev = static_cast<int64_t*>(g->scratch("hwe_rows", g->n * sizeof(int64_t)));
                GPE_HIP(hipMemcpyAsync(rows_dev, global_rows_host, g->n * sizeof(int64_t), hipMemcpyHostToDevice,
                                       st));
            }
            launch_hwe_impute(g->snp_work, g->snp_stride, g->m, g->n, g->nobs, g->sum, impute_seed, snp_offset,
                              rows_dev, st);
            launch_snp_stats(g->snp_work, g->snp_stride, g->m, g->n, g->nobs, g->sum, g->sumsq, st);
            launch_std_params(g->nobs, g->sum, g->sumsq, g->m, g->n, ddof, 1, g->mean, g->inv_sd, st);
        } else {
            launch_std_params(g->nobs, g->sum, g->sumsq, g->m, g->n, ddof, 0, g->mean, g->inv_sd, st);
        }
        launch_transpose(g->snp_work, g->snp_stride, g->m, g->smp_work, g->smp_stride, g->n, st);
        // per-sample missing-call lists (mean mode only: HWE leaves none)
        int32_t* counts = static_cast<int32_t*>(g->scratch("miss_counts", g->n * sizeof(int32_t)));
        launch_miss_count(g->smp_work, g->smp_stride, g->n, counts, st);
        g->miss_ptr = dalloc<int64_t>(g->n + 1);
        launch_scan_counts(counts, g->n, g->miss_ptr, st);
        int64_t total = 0;
        GPE_HIP(hipMemcpyAsync(&total, g->miss_ptr + g->n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
        GPE_HIP(hipStreamSynchronize(st));
        g->n_missing = total;
        if (total > 0) {
            g->miss_idx = dalloc<int32_t>(total);
            launch_miss_fill(g->smp_work, g->smp_stride, g->n, g->miss_ptr, g->miss_idx, st);
            // the same calls listed per SNP (phase-2 missing-call sums)
            int32_t* scounts = static_cast<int32_t*>(g->scratch("smiss_counts", g->m * sizeof(int32_t)));
            launch_miss_count(g->snp_work, g->snp_stride, g->m, scounts, st);
            g->smiss_ptr = dalloc<int64_t>(g->m + 1);
            launch_scan_counts(scounts, g->m, g->smiss_ptr, st);
            int64_t total_snp = 0;
            GPE_HIP(hipMemcpyAsync(&total_snp, g->smiss_ptr + g->m, sizeof(int64_t), hipMemcpyDeviceToHost, st));
            GPE_HIP(hipStreamSynchronize(st));
            require(total_snp == total, "missing-call lists by SNP and by sample disagree");
            g->smiss_idx = dalloc<int32_t>(total);
            launch_miss_fill(g->snp_work, g->snp_stride, g->m, g->smiss_ptr, g->smiss_idx, st);
        }
        // the contractions stream chunk-tiled copies (one step's 32 rows x 32 bytes are
        // 1 KiB of consecutive bytes instead of 32 pieces of 32 bytes, one per row) when
        // both copies fit with a quarter of the device's memory to spare; otherwise they
        // read the row-major layouts (the same results).  SUMMIT_AMD_TILED=0 turns the
        // copies off.
        const char* tiled_env = std::getenv("SUMMIT_AMD_TILED");
        const bool want_tiled = !(tiled_env && tiled_env[0] == '0');
        const size_t tiled_bytes = (size_t)(g->m_alloc * g->snp_stride + g->n_alloc * g->smp_stride);
        size_t free_b = 0, total_b = 0;
        GPE_HIP(hipStreamSynchronize(st));
        GPE_HIP(hipMemGetInfo(&free_b, &total_b));
        if (want_tiled && tiled_bytes + total_b / 4 <= free_b) {
            g->snp_tiled = dalloc<uint8_t>(g->m_alloc * g->snp_stride);
            launch_tile_packed(g->snp_work, g->snp_stride, g->m_alloc, g->snp_tiled, st);
            g->smp_tiled = dalloc<uint8_t>(g->n_alloc * g->smp_stride);
            launch_tile_packed(g->smp_work, g->smp_stride, g->n_alloc, g->smp_tiled, st);
        }
        GPE_HIP(hipStreamSynchronize(st));
        g->prepared = true;
    });
}

int gpe_geno_stats(gpe_geno* g, int32_t* nobs_host, int64_t* sum_host, int64_t* sumsq_host, double* mean_host,
                   double* inv_sd_host) {
    return guarded([&] {
        require_prepared(g);
        if (nobs_host) GPE_HIP(hipMemcpy(nobs_host, g->nobs, g->m * sizeof(int32_t), hipMemcpyDeviceToHost));
        if (sum_host) GPE_HIP(hipMemcpy(sum_host, g->sum, g->m * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (sumsq_host) GPE_HIP(hipMemcpy(sumsq_host, g->sumsq, g->m * sizeof(int64_t), hipMemcpyDeviceToHost));
        if (mean_host) GPE_HIP(hipMemcpy(mean_host, g->mean, g->m * sizeof(double), hipMemcpyDeviceToHost));
        if (inv_sd_host) GPE_HIP(hipMemcpy(inv_sd_host, g->inv_sd, g->m * sizeof(double), hipMemcpyDeviceToHost));
    });
}

int gpe_geno_download_rows(gpe_geno* g, int32_t layout, int64_t row_begin, int64_t row_count, uint8_t* host,
                           int64_t host_stride) {
    return guarded([&] {
        require(g != nullptr && host != nullptr, "null argument to gpe_geno_download_rows");
        const bool snp = layout == GPE_LAYOUT_SNP_RAW || layout == GPE_LAYOUT_SNP_WORK;
        require(snp || (layout == GPE_LAYOUT_SAMPLE_WORK && g->prepared), "unknown or unprepared layout");
        const int64_t rows = snp ? g->m : g->n;
        const int64_t stride = snp ? g->snp_stride : g->smp_stride;
        const int64_t used = snp ? ceil_div(g->n, 4) : ceil_div(g->m, 4);
        const uint8_t* src = layout == GPE_LAYOUT_SNP_RAW ? g->snp_raw : (snp ? g->snp_work : g->smp_work);
        require(row_begin >= 0 && row_count >= 0 && row_begin + row_count <= rows, "row range out of bounds");
        require(host_stride >= used, "host_stride too small");
        if (row_count == 0) return;
        download_rows_staged(g, host, host_stride, src + row_begin * stride, stride, used, row_count);
    });
}

int gpe_geno_download(gpe_geno* g, int32_t layout, uint8_t* host, int64_t host_stride) {
    return guarded([&] {
        require(g != nullptr && host != nullptr, "null argument to gpe_geno_download");
        const uint8_t* src;
        int64_t rows, stride, used;
        if (layout == GPE_LAYOUT_SNP_RAW || layout == GPE_LAYOUT_SNP_WORK) {
            src = layout == GPE_LAYOUT_SNP_RAW ? g->snp_raw : g->snp_work;
            rows = g->m;
            stride = g->snp_stride;
            used = ceil_div(g->n, 4);
        } else if (layout == GPE_LAYOUT_SAMPLE_WORK) {
            require(g->prepared, "sample-major layout exists after gpe_geno_prepare");
            src = g->smp_work;
            rows = g->n;
            stride = g->smp_stride;
            used = ceil_div(g->m, 4);
        } else {
            throw Error("unknown layout");
        }
        require(host_stride >= used, "host_stride too small");
        download_rows_staged(g, host, host_stride, src, stride, used, rows);
    });
}

int gpe_resvar_inv(gpe_geno* g, int64_t snp_begin, int64_t snp_end, double eps, double* inv_out_dev,
                   void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "SNP range out of bounds");
        require(g->n > 1, "Invalid block dimensions in compute_inv_resvar_from_geno_block");
        launch_resvar_inv(g->nobs, g->sum, g->sumsq, g->mean, g->inv_sd, snp_begin, snp_end, g->n,
                          g->impute_mode == GPE_IMPUTE_HWE ? 1 : 0, eps, inv_out_dev, S(stream));
    });
}

int gpe_gz_i32(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int8_t* z_dev, int64_t ldz, int32_t n_cols,
               int32_t* u_dev, int64_t ldu, int32_t accumulate, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin < snp_end, "SNP range out of bounds");
        require(n_cols > 0 && ldz >= n_cols && ldu >= g->n, "bad leading dimension");
        const int64_t L = snp_end - snp_begin;
        require(L * 2 * 128 < ((int64_t)1 << 31), "too many SNPs for one int32 product; split the range");
        hipStream_t st = S(stream);
        const int ctc = 2;
        const int groups = (int)ceil_div(n_cols, 32 * ctc);
        const int64_t coef_ctiles = (int64_t)groups * ctc;
        const int64_t kc0 = snp_begin / kKChunk, kc1 = ceil_div(snp_end, kKChunk);
        const int64_t kcount = kc1 - kc0;
        int8_t* coef = static_cast<int8_t*>(g->scratch("coef", kcount * coef_ctiles * 4096));
        launch_coef_i8(z_dev, ldz, snp_begin, snp_end, n_cols, coef_ctiles, coef, st);
        if (!accumulate)
            for (int c = 0; c < n_cols; ++c)
                launch_fill_u8(reinterpret_cast<uint8_t*>(u_dev + (int64_t)c * ldu), g->n * 4, 0, st);
        ContractArgs a{};
        a.packed = g->smp_tiled ? g->smp_tiled : g->smp_work;
        a.tiled = g->smp_tiled ? 1 : 0;
        a.stride = g->smp_stride;
        a.r0 = 0;
        a.r1 = g->n;
        a.coef = coef;
        a.coef_ctiles = coef_ctiles;
        a.kc0 = kc0;
        a.kc1 = kc1;
        a.ctiles_used = (int)ceil_div(n_cols, 32);
        a.out_i32 = u_dev;
        a.ldo = ldu;
        a.cols_out = n_cols;
        a.out_row0 = 0;
        const int64_t gx = ceil_div(g->n, 256);
        const int splits = pick_splits(gx * groups, kcount, 1 << 30, contract_slots(1, ctc, false, EPI_I32_ATOMIC));
        a.kc_per_split = ceil_div(kcount, splits);
        ProfScope ps(GPE_PROF_GZ, st);
        // two 128-value chunks per LDS slot: a sample row's two 32-byte pieces are
        // loaded together, so a 64-byte run of its L2 line is used at once
        // (N=100,000, M=500,000, 64 columns: 3.88 ms against 4.23 ms with one chunk)
        launch_contract(1, ctc, false, EPI_I32_ATOMIC, a, groups, splits, st, 2);
        ps.stop();
    });
}

int gpe_gtu_i64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int32_t* u_dev, int64_t ldu,
                int32_t n_cols, int64_t* w_dev, int64_t ldw, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin < snp_end, "SNP range out of bounds");
        require(n_cols > 0 && ldu >= g->n && ldw >= n_cols, "bad leading dimension");
        hipStream_t st = S(stream);
        // |U| must fit three balanced int8 limbs
        unsigned int* mx = static_cast<unsigned int*>(g->scratch("absmax", sizeof(unsigned int)));
        launch_fill_u8(reinterpret_cast<uint8_t*>(mx), 4, 0, st);
        hipLaunchKernelGGL(k_absmax_i32, dim3(64, (unsigned)n_cols), dim3(256), 0, st, u_dev, ldu, g->n, n_cols, mx);
        GPE_HIP(hipGetLastError());
        unsigned int mx_h = 0;
        GPE_HIP(hipMemcpyAsync(&mx_h, mx, 4, hipMemcpyDeviceToHost, st));
        GPE_HIP(hipStreamSynchronize(st));
        require(mx_h <= 8421504u, "|U| exceeds the three-limb range (8421504)");
        const int nl = 3, ctc = 1;
        const int groups = (int)ceil_div(n_cols, 32 * ctc);
        const int64_t coef_ctiles = (int64_t)groups * nl * ctc;
        const int64_t kcount = ceil_div(g->n, kKChunk);
        int8_t* coef = static_cast<int8_t*>(g->scratch("coef", kcount * coef_ctiles * 4096));
        launch_quant_i32(u_dev, ldu, g->n, n_cols, nl, ctc, coef_ctiles, coef, st);
        ContractArgs a{};
        a.packed = g->snp_tiled ? g->snp_tiled : g->snp_work;
        a.tiled = g->snp_tiled ? 1 : 0;
        a.stride = g->snp_stride;
        a.r0 = snp_begin;
        a.r1 = snp_end;
        a.coef = coef;
        a.coef_ctiles = coef_ctiles;
        a.kc0 = 0;
        a.kc1 = kcount;
        a.kc_per_split = kcount;
        a.ctiles_used = (int)ceil_div(n_cols, 32);
        a.out_i64 = w_dev;
        a.ldo = ldw;
        a.cols_out = n_cols;
        a.out_row0 = snp_begin;
        ProfScope ps(GPE_PROF_GTU, st);
        // two chunks per LDS slot for the 3-limb product (7.55 ms against 7.76 ms at the
        // same shape)
        launch_contract(nl, ctc, false, EPI_I64_ROWMAJOR, a, groups, 1, st, nl == 3 ? 2 : 1);
        ps.stop();
    });
}

int gpe_probes(int32_t rand_dist, uint64_t root_seed, const int64_t* blk_starts_host, int32_t n_blocks,
               int64_t blk_end, int32_t v_start, int32_t v_count, double* z_host) {
    return guarded([&] {
        require(n_blocks > 0 && v_count > 0 && blk_starts_host && z_host, "bad probe request");
        const int64_t s0 = blk_starts_host[0];
        const int64_t L = blk_end - s0;
        require(L > 0, "empty probe range");
        if (rand_dist == GPE_RAND_RADEMACHER) {
            int64_t* starts = dalloc<int64_t>(n_blocks);
            int8_t* z = dalloc<int8_t>((s0 + L) * (int64_t)v_count);
            std::vector<int8_t> zh((size_t)(L * v_count));
            try {
                GPE_HIP(hipMemcpy(starts, blk_starts_host, n_blocks * sizeof(int64_t), hipMemcpyHostToDevice));
                launch_mt_rademacher(root_seed, starts, n_blocks, blk_end, v_start, v_count, z, v_count, nullptr);
                GPE_HIP(hipDeviceSynchronize());
                GPE_HIP(hipMemcpy(zh.data(), z + s0 * v_count, zh.size(), hipMemcpyDeviceToHost));
            } catch (...) {
                dfree(starts);
                dfree(z);
                throw;
            }
            dfree(starts);
            dfree(z);
            for (size_t i = 0; i < zh.size(); ++i) z_host[i] = (double)zh[i];
        } else if (is_real_probe_dist(rand_dist)) {
            real_probes_into(rand_dist, root_seed, blk_starts_host, n_blocks, blk_end, v_start, v_count, z_host,
                             v_count);
        } else {
            throw Error("rand_dist must be rademacher, gaussian or spherical");
        }
    });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Shared drivers of the two traversals
// ---------------------------------------------------------------------------
namespace {

// Phase-1 type product on the sample-major layout:
//   out[(bin*v_count + v)][n] (=|+=) sum_{i in [s, e)} (g_in - mu_i m_in) * w[i][bin] * z[i][v]
// (x_in = (g_in - mu_i) inv_sd_i m_in when w carries inv_sd).  w, z indexed by absolute SNP.
void run_phase1(gpe_geno* g, int64_t s, int64_t e, const double* w, int n_bins, const int8_t* z_i8,
                const double* z_f64, int64_t ldz, int v_count, double* out, int64_t ldo, const uint8_t* bin_mode,
                int limbs, hipStream_t st) {
    const LimbCfg lc = limb_cfg(limbs);
    const int64_t L = e - s;
    const int v_pad = (int)round_up(v_count, 32);
    require(ldz >= v_pad, "probe leading dimension too small");
    require((int64_t)n_bins * v_pad <= 65535,
            "bins x padded probes exceeds 65535 columns per traversal; split the probes into V-tiles");
    const int ncols = n_bins * v_pad;
    const int groups = (int)ceil_div(ncols, 32 * lc.ctc);
    const int64_t coef_ctiles = (int64_t)groups * lc.nl * lc.ctc;
    const int ncols_pad = groups * lc.ctc * 32;
    uint64_t* colmax = static_cast<uint64_t*>(g->scratch("colmax", ncols_pad * sizeof(uint64_t)));
    int32_t* shift = static_cast<int32_t*>(g->scratch("shift", ncols_pad * sizeof(int32_t)));
    double* col_scale = static_cast<double*>(g->scratch("col_scale", ncols_pad * sizeof(double)));
    double* cst = static_cast<double*>(g->scratch("cst", ncols_pad * sizeof(double)));
    launch_fill_u64(colmax, ncols_pad, 0, st);
    launch_coef_p1_colmax(w, n_bins, z_i8, z_f64, ldz, s, e, n_bins, v_count, v_pad, colmax, st);
    const int prec = std::min(8 * lc.nl - 2, 62 - bitlen((uint64_t)(2 * L)));
    launch_shift_from_colmax(colmax, ncols_pad, prec, shift, col_scale, st);
    const int64_t kc0 = s / kKChunk, kc1 = ceil_div(e, kKChunk);
    const int64_t kcount = kc1 - kc0;
    int8_t* coef = static_cast<int8_t*>(g->scratch("coef", kcount * coef_ctiles * 4096));
    // the quantizer writes every (value, column) of the first ncols columns; only
    // padding column tiles need clearing
    if (ncols_pad != ncols) launch_fill_u8(reinterpret_cast<uint8_t*>(coef), kcount * coef_ctiles * 4096, 0, st);
    double* cst_part = static_cast<double*>(g->scratch("cst_part", (size_t)ncols_pad * kCstParts * sizeof(double)));
    launch_coef_p1_quant(w, n_bins, z_i8, z_f64, ldz, s, e, n_bins, v_count, v_pad, shift, lc.nl, lc.ctc,
                         coef_ctiles, kc0, coef, g->mean, cst, cst_part, st);
    const int64_t ldy = g->n_alloc;
    int64_t* y = static_cast<int64_t*>(g->scratch("p1_y", (int64_t)ncols_pad * ldy * sizeof(int64_t)));
    launch_fill_u64(reinterpret_cast<uint64_t*>(y), (int64_t)ncols_pad * ldy, 0, st);
    ContractArgs a{};
    a.packed = g->smp_tiled ? g->smp_tiled : g->smp_work;
    a.tiled = g->smp_tiled ? 1 : 0;
    a.stride = g->smp_stride;
    a.r0 = 0;
    a.r1 = g->n;
    a.coef = coef;
    a.coef_ctiles = coef_ctiles;
    a.kc0 = kc0;
    a.kc1 = kc1;
    a.ctiles_used = ncols / 32;
    a.out_i64 = y;
    a.ldo = ldy;
    a.cols_out = ncols_pad;
    a.out_row0 = 0;
    const int64_t gx = ceil_div(g->n, 256);
    // int32 per-limb partial sums over a split's K range: at most 2^31 / (256 * 128) chunks per split
    const int64_t max_chunks = (int64_t)(((int64_t)1 << 31) / (256 * 128)) - 1;
    int splits = sched_lookup(GPE_SCHED_P1_SPLITS, g->n, L, ncols_pad, lc.nl);
    if (splits <= 0)
        splits = pick_splits_deep(gx * groups, kcount, max_chunks, contract_slots(lc.nl, lc.ctc, false, EPI_I64_ATOMIC));
    splits = (int)std::min<int64_t>(std::max<int64_t>(splits, ceil_div(kcount, max_chunks)), kcount);
    a.kc_per_split = ceil_div(kcount, splits);
    splits = (int)ceil_div(kcount, a.kc_per_split);
    // two chunks per barrier unless the schedule says one (measured ~3% faster at the headline shape)
    const int cpb = sched_lookup(GPE_SCHED_P1_CPB, g->n, L, ncols_pad, lc.nl) == 1 ? 1 : 2;
    ProfScope ps(GPE_PROF_P1_CONTRACT, st);
    launch_contract(lc.nl, lc.ctc, false, EPI_I64_ATOMIC, a, groups, splits, st, cpb);
    ps.stop();
    const bool miss = g->impute_mode == GPE_IMPUTE_MEAN && g->n_missing > 0;
    // missing calls with int8 probes: per-call records summed by k_p1_finalize_z8, or the
    // missing-flag plane as a second traversal (always for real-valued probes).  Per sample
    // and 64-probe word the records cost about 3.48 ns + 12.2 ps per missing call, the plane
    // (quantization, traversal, combine) about 0.96 ns + 0.248 ps per SNP -- fitted to
    // N=100,000, M=500,000, 64 probes, 0.5% missing (records 3.4 ms, plane 12.5 ms) and
    // N=5,000, M=10,000, 8,192 probes, 1% missing (records 3.0 ms, plane 2.2 ms) -- so the
    // plane wins when L (0.248 ps - 12.2 ps d) < 2.52 ns for L SNPs at missing density d
    const int p1m = sched_lookup(GPE_SCHED_P1_MISS, g->n, L, ncols_pad, lc.nl);
    const double density = (double)g->n_missing / ((double)g->n * (double)g->m);
    const bool plane_cheaper = (double)L * (0.248e-12 - 12.2e-12 * density) < 2.52e-9;
    const bool plane = miss && (z_f64 != nullptr || p1m == 2 || (p1m != 1 && plane_cheaper));
    if (miss && z_i8 && !plane) {
        void* rec = g->scratch("p1_mrec", p1_miss_rec_bytes(L, n_bins, v_pad));
        launch_p1_miss_rec(w, n_bins, z_i8, ldz, s, e, n_bins, v_count, v_pad, shift, col_scale, g->mean, rec, st);
        const int32_t kib = sched_lookup(GPE_SCHED_P1_FINALIZE_KIB, g->n, L, ncols_pad, lc.nl);
        ProfScope pf(GPE_PROF_P1_FINALIZE, st);
        launch_p1_finalize_z8(y, ldy, g->n, n_bins, v_count, v_pad, col_scale, cst, g->miss_ptr, g->miss_idx, s, e,
                              rec, out, ldo, bin_mode, (int64_t)kib << 10, st);
        pf.stop();
        return;
    }
    if (plane) {
        // the missing-call term sum_{i missing at n} mu_i c_iv as a second int8-MFMA traversal:
        // the missing-flag plane against the quantized mu_i w_ik z_iv (for real-valued probes a
        // per-call record would carry 64 FP64 values per call)
        ProfScope pf(GPE_PROF_P1_FINALIZE, st);
        double* w2 = static_cast<double*>(g->scratch("p1_w2", g->m_alloc * n_bins * sizeof(double)));
        launch_scale_rows(w, g->mean, s, e, n_bins, w2, st);
        uint64_t* colmax2 = static_cast<uint64_t*>(g->scratch("colmax2", ncols_pad * sizeof(uint64_t)));
        int32_t* shift2 = static_cast<int32_t*>(g->scratch("shift2", ncols_pad * sizeof(int32_t)));
        double* cs2 = static_cast<double*>(g->scratch("col_scale2", ncols_pad * sizeof(double)));
        launch_fill_u64(colmax2, ncols_pad, 0, st);
        launch_coef_p1_colmax(w2, n_bins, z_i8, z_f64, ldz, s, e, n_bins, v_count, v_pad, colmax2, st);
        launch_shift_from_colmax(colmax2, ncols_pad, prec, shift2, cs2, st);
        // the main contraction has consumed the coefficient buffer (stream order); reuse it
        if (ncols_pad != ncols) launch_fill_u8(reinterpret_cast<uint8_t*>(coef), kcount * coef_ctiles * 4096, 0, st);
        // (no Cst: the flag plane has no mean-centring term)
        launch_coef_p1_quant(w2, n_bins, z_i8, z_f64, ldz, s, e, n_bins, v_count, v_pad, shift2, lc.nl, lc.ctc,
                             coef_ctiles, kc0, coef, g->mean, nullptr, nullptr, st);
        int64_t* y2 = static_cast<int64_t*>(g->scratch("p1_y2", (int64_t)ncols_pad * ldy * sizeof(int64_t)));
        launch_fill_u64(reinterpret_cast<uint64_t*>(y2), (int64_t)ncols_pad * ldy, 0, st);
        ContractArgs a2 = a;
        a2.out_i64 = y2;
        launch_contract(lc.nl, lc.ctc, false, EPI_I64_ATOMIC_Q, a2, groups, splits, st, cpb);
        launch_p1_combine2(y, y2, ldy, g->n, n_bins, v_count, v_pad, col_scale, cst, cs2, out, ldo, bin_mode, st);
        pf.stop();
        return;
    }
    launch_p1_finalize(y, ldy, g->n, n_bins, v_count, v_pad, col_scale, cst, miss ? g->miss_ptr : nullptr,
                       miss ? g->miss_idx : nullptr, s, e, w, n_bins, z_i8, z_f64, ldz, shift, g->mean, out, ldo,
                       bin_mode, st);
}

// Missing-call density (missing calls / (N M)) up to which phase 2 gathers the
// missing-call sums instead of running a missing-flag MFMA plane: the gathers
// cost about 7 ms per 0.5% at N = 100,000, M = 500,000, 64 columns, the plane
// about 8 ms at any density (break-even near 0.55%)
constexpr double kSparseMissDensity = 0.0055;
// sample ranges of the gathered missing-call sums (one per XCD)
constexpr int kMissParts = 8;

// Phase-2 type product on the SNP-major layout against x (column-major (N, n_bins*v_count)):
// quantize x into limbs and run the contraction with the given epilogue.
// sum_override (may be null): per column of x, sums that replace the exact column
// sums of the quantized operand in the centring term.
void run_phase2(gpe_geno* g, int64_t s, int64_t e, const double* x, int64_t ldx, int n_bins, int v_count, int limbs,
                EpiMode mode, ContractArgs& a, hipStream_t st, int* ctiles_used_out, int* v_pad_out,
                const double* sum_override = nullptr) {
    const LimbCfg lc = limb_cfg(limbs);
    const int v_pad = (int)round_up(v_count, 32);
    require((int64_t)n_bins * v_pad <= 65535,
            "bins x padded probes exceeds 65535 columns per traversal; split the probes into V-tiles");
    const int ncols = n_bins * v_pad;
    const int groups = (int)ceil_div(ncols, 32 * lc.ctc);
    const int64_t coef_ctiles = (int64_t)groups * lc.nl * lc.ctc;
    const int ncols_pad = groups * lc.ctc * 32;
    const int64_t kcount = ceil_div(g->n, kKChunk);
    uint64_t* colmax = static_cast<uint64_t*>(g->scratch("colmax", ncols_pad * sizeof(uint64_t)));
    int32_t* shift = static_cast<int32_t*>(g->scratch("shift", ncols_pad * sizeof(int32_t)));
    double* col_scale = static_cast<double*>(g->scratch("col_scale", ncols_pad * sizeof(double)));
    int64_t* col_sum = static_cast<int64_t*>(g->scratch("col_sum", ncols_pad * sizeof(int64_t)));
    launch_fill_u64(colmax, ncols_pad, 0, st);
    launch_fill_u64(reinterpret_cast<uint64_t*>(col_sum), ncols_pad, 0, st);
    launch_colmax_f64(x, ldx, g->n, ncols, v_count, v_pad, colmax, st);
    const int prec = std::min(8 * lc.nl - 2, 62 - bitlen((uint64_t)(2 * g->n)));
    launch_shift_from_colmax(colmax, ncols_pad, prec, shift, col_scale, st);
    int8_t* coef = static_cast<int8_t*>(g->scratch("coef", kcount * coef_ctiles * 4096));
    if (ncols_pad != ncols) launch_fill_u8(reinterpret_cast<uint8_t*>(coef), kcount * coef_ctiles * 4096, 0, st);
    launch_quant_p2(x, ldx, g->n, n_bins, v_count, v_pad, shift, lc.nl, lc.ctc, coef_ctiles, coef, col_sum, st);
    if (sum_override) launch_col_sum_override(sum_override, n_bins, v_count, v_pad, shift, col_scale, col_sum, st);
    a.packed = g->snp_tiled ? g->snp_tiled : g->snp_work;
    a.tiled = g->snp_tiled ? 1 : 0;
    a.stride = g->snp_stride;
    a.r0 = s;
    a.r1 = e;
    a.coef = coef;
    a.coef_ctiles = coef_ctiles;
    a.kc0 = 0;
    a.kc1 = kcount;
    a.kc_per_split = kcount;
    a.ctiles_used = ncols / 32;
    a.mean = g->mean;
    a.inv_sd = g->inv_sd;
    a.col_scale = col_scale;
    a.col_sum = col_sum;
    const bool miss = g->impute_mode == GPE_IMPUTE_MEAN && g->n_missing > 0;
    a.has_miss = miss ? 1 : 0;
    const int cpb = sched_lookup(GPE_SCHED_P2_CPB, e - s, g->n, ncols_pad, lc.nl) == 1 ? 1 : 2;
    // missing calls: a second (missing-flag) MFMA plane, or exact sums gathered over
    // the per-SNP missing lists when calls are sparse (the same int64 values)
    const int p2m = sched_lookup(GPE_SCHED_P2_MISS, e - s, g->n, ncols_pad, lc.nl);
    const double density = (double)g->n_missing / ((double)g->n * (double)g->m);
    const bool gather = miss && mode == EPI_P2 && (p2m == 2 || (p2m != 1 && density <= kSparseMissDensity));
    if (gather) {
        const int nch = (int)ceil_div(ncols_pad, 64);
        int parts = sched_lookup(GPE_SCHED_P2_MISS_PARTS, e - s, g->n, ncols_pad, lc.nl);
        if (parts <= 0) parts = kMissParts;
        // per-SNP list offsets of the sample ranges, kept on the handle for the range count in use
        if (g->rptr_parts != parts) {
            require(!g->capturing, "missing-list ranges not on the device: run the same call once outside graph "
                                   "capture first");
            dfree(g->smiss_rptr);
            g->smiss_rptr = nullptr;
            g->rptr_parts = 0;
            g->smiss_rptr = dalloc<int64_t>(g->m * (parts + 1));
            launch_range_ptrs(g->smiss_ptr, g->smiss_idx, g->m, parts, ceil_div(g->n, parts), g->smiss_rptr, st);
            g->rptr_parts = parts;
        }
        int8_t* xrow = static_cast<int8_t*>(g->scratch("p2_xrow", (size_t)(g->n + 1) * nch * lc.nl * 64));
        launch_xrow_limbs(x, ldx, g->n, n_bins, v_count, v_pad, shift, lc.nl, nch, xrow, st);
        const int64_t ld_ms = (int64_t)nch * 64;
        // SNP slices keep the sums near 4 GiB and the gather grid within 2^25 waves
        // (2^31 threads in x)
        const int64_t by_bytes = round_up(((int64_t)1 << 32) / (ld_ms * 8), 256);
        const int64_t by_waves = (((int64_t)1 << 25) / ((int64_t)parts * nch)) / 256 * 256;
        const int64_t slice = std::max<int64_t>(256, std::min(by_bytes, by_waves));
        const int64_t rows = std::min<int64_t>(slice, e - s);
        int64_t* msum = static_cast<int64_t*>(g->scratch("p2_msum", (size_t)rows * ld_ms * sizeof(int64_t)));
        a.msum = msum;
        a.ld_ms = ld_ms;
        for (int64_t rs = s; rs < e; rs += slice) {
            const int64_t re = std::min(e, rs + slice);
            ProfScope pm(GPE_PROF_P2_MISS, st);
            // the sample ranges' sums meet in one zeroed buffer (the epilogue reads one value)
            launch_fill_u64(reinterpret_cast<uint64_t*>(msum), (re - rs) * ld_ms, 0, st);
            launch_miss_sums(g->smiss_rptr, parts, g->smiss_idx, rs, re, xrow, g->n, lc.nl, nch, msum, ld_ms, st);
            pm.stop();
            a.r0 = rs;
            a.r1 = re;
            ProfScope ps(GPE_PROF_P2_CONTRACT, st);
            launch_contract(lc.nl, lc.ctc, false, EPI_P2X, a, groups, 1, st, cpb);
            ps.stop();
        }
        a.r0 = s;
        a.r1 = e;
        a.msum = nullptr;
        if (ctiles_used_out) *ctiles_used_out = ncols / 32;
        if (v_pad_out) *v_pad_out = v_pad;
        return;
    }
    ProfScope ps(GPE_PROF_P2_CONTRACT, st);
    launch_contract(lc.nl, lc.ctc, miss, mode, a, groups, 1, st, cpb);
    ps.stop();
    if (ctiles_used_out) *ctiles_used_out = ncols / 32;
    if (v_pad_out) *v_pad_out = v_pad;
}

// W_sm[s][e] = w[e][s] for s < n and e < n_w, 0 elsewhere: (n_pad, ec) sample-major
__global__ void k_weights_sample_major(const double* w, int64_t ldw, int n_w, int64_t n, int64_t n_pad, int ec,
                                       double* wsm) {
    const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= n_pad * ec) return;
    const int64_t s = idx / ec;
    const int e = (int)(idx % ec);
    wsm[idx] = (s < n && e < n_w) ? w[(int64_t)e * ldw + s] : 0.0;
}

__global__ void k_w_from_inv_sd(const double* inv_sd, const int32_t* nobs, const int64_t* sum, const int64_t* sumsq,
                                int64_t s, int64_t e, double* w) {
    const int64_t i = s + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < e) w[i] = zero_column(nobs, sum, sumsq, i) ? 0.0 : inv_sd[i];
}

// copy a row-major (rows, cols) matrix into rows [row0, row0+rows) of a buffer with leading dimension ld
void copy_rows(const double* src, int64_t lds, int64_t rows, int cols, double* dst, int64_t ld, int64_t row0,
               hipStream_t st) {
    GPE_HIP(hipMemcpy2DAsync(dst + row0 * ld, ld * sizeof(double), src, lds * sizeof(double), cols * sizeof(double),
                             rows, hipMemcpyDeviceToDevice, st));
}

// Exact windowed LD of n_chr chromosomes [bounds[2c], bounds[2c+1]) (sorted,
// non-overlapping, absolute SNP ids) in ONE banded-Gram launch: every
// chromosome's 64x64 tile pairs within the window go into one work list.
// Rows of bp (when used), annot and out are SNP ids relative to bounds[0].
// Covariates (p > 0): C column-major (N, p) with ldc, R row-major (p, N) with ldr;
// every panel is projected, X <- X - C (R X), and restandardized
// (src/native/winldcore.cpp:341-372), corr_dim = N - p (:848-851).  The projection
// enters the exact integer Gram as a rank-2p FP64 correction of each pair's cross
// product plus per-SNP means and scales (k_win_cov_prep).
void windowed_ld_impl(gpe_geno* g, int n_chr, const int64_t* bounds, const int64_t* bp_host, double window_bp,
                      int64_t window_snps, const double* annot_dev, int32_t n_bins, double* out_dev,
                      hipStream_t st, const double* c_dev = nullptr, int64_t ldc = 0, const double* r_dev = nullptr,
                      int64_t ldr = 0, int p = 0, int side = 0) {
    require(n_bins > 0, "annot_chr must have at least one column");
    require(g->n > 2, "Too few rows for windowed LD score computation");
    require(p >= 0 && p <= 1024, "covariate count out of range");
    if (p > 0)
        require(c_dev != nullptr && r_dev != nullptr && ldc >= g->n && ldr >= g->n,
                "C/R shape mismatch in compute_windowed_ld_chr");
    const double corr_dim = (double)g->n - (p > 0 ? (double)p : 1.0);
    require(corr_dim > 1.0, "Too few residual dimensions for unbiased windowed r^2 correction");
    const int64_t row0 = bounds[0];
    for (int c = 0; c < n_chr; ++c) {
        const int64_t b = bounds[2 * c], e = bounds[2 * c + 1];
        require(e > b, "Chromosome slice is empty");
        require(b >= 0 && e <= g->m, "SNP range out of bounds");
        require(c == 0 || b >= bounds[2 * c - 1], "chromosome SNP ranges must be sorted and must not overlap");
    }
    const int64_t rows = bounds[2 * n_chr - 1] - row0;
    if (window_snps <= 0) {
        require(bp_host != nullptr, "bp length mismatch in compute_windowed_ld_chr");
        require(std::isfinite(window_bp) && window_bp > 0.0, "ld_wind_kb must define a finite positive BP window");
        for (int c = 0; c < n_chr; ++c)
            for (int64_t i = bounds[2 * c] + 1; i < bounds[2 * c + 1]; ++i)
                require(bp_host[i - row0] >= bp_host[i - 1 - row0],
                        "BP not sorted within chromosome block. Sort your .bim by CHR+BP.");
    }
    // (ti, tj, chromosome) tile pairs of every band
    std::vector<int32_t> pairs;
    for (int c = 0; c < n_chr; ++c) {
        const int64_t b = bounds[2 * c], m = bounds[2 * c + 1] - b;
        const int64_t nt = ceil_div(m, 64);
        for (int64_t ti = 0; ti < nt; ++ti) {
            const int64_t i_last = std::min(m, (ti + 1) * 64) - 1;
            for (int64_t tj = ti; tj < nt; ++tj) {
                const int64_t j_first = tj * 64;
                if (tj > ti) {
                    const bool far = window_snps > 0
                                         ? (j_first - i_last > window_snps)
                                         : ((double)(bp_host[b - row0 + j_first] - bp_host[b - row0 + i_last]) >
                                            window_bp);
                    if (far) break;
                }
                pairs.push_back((int32_t)ti);
                pairs.push_back((int32_t)tj);
                pairs.push_back((int32_t)c);
            }
        }
    }
    const int64_t np = (int64_t)pairs.size() / 3;
    int32_t* pairs_dev = static_cast<int32_t*>(g->scratch("win_pairs", pairs.size() * sizeof(int32_t)));
    GPE_HIP(hipMemcpyAsync(pairs_dev, pairs.data(), pairs.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    int64_t* chr_dev = static_cast<int64_t*>(g->scratch("win_chr", 2 * (size_t)n_chr * sizeof(int64_t)));
    GPE_HIP(hipMemcpyAsync(chr_dev, bounds, 2 * (size_t)n_chr * sizeof(int64_t), hipMemcpyHostToDevice, st));
    int64_t* bp_dev = nullptr;
    if (window_snps <= 0) {
        bp_dev = static_cast<int64_t*>(g->scratch("win_bp", rows * sizeof(int64_t)));
        GPE_HIP(hipMemcpyAsync(bp_dev, bp_host, rows * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    double* scale = static_cast<double*>(g->scratch("win_scale", g->m_alloc * sizeof(double)));
    const int completed = g->impute_mode == GPE_IMPUTE_HWE ? 1 : 0;
    launch_window_scale(g->nobs, g->sum, g->sumsq, g->m, g->n, completed, scale, st);
    launch_fill_u8(reinterpret_cast<uint8_t*>(out_dev), rows * n_bins * (int64_t)sizeof(double), 0, st);
    GramArgs a{};
    a.packed = g->snp_work;
    a.stride = g->snp_stride;
    a.n = g->n;
    a.kc1 = ceil_div(g->n, kKChunk);
    a.pairs = pairs_dev;
    a.chr = chr_dev;
    a.row0 = row0;
    a.mean = g->mean;
    a.scale = scale;
    a.sum = g->sum;
    a.nobs = g->nobs;
    a.bp = bp_dev;
    a.window_bp = window_bp;
    a.window_snps = window_snps;
    a.annot = annot_dev;
    a.n_bins = n_bins;
    a.out = out_dev;
    a.corr_dim = corr_dim;
    require(side >= 0 && side <= 2, "window side must be 0, 1 or 2");
    a.side = side;
    if (p > 0) {
        // u = C^T x_j and v = R x_j for every SNP of the rows (exact-limb FP64 products),
        // then the per-SNP correction rows, means and scales of the projected columns
        double* u = static_cast<double*>(g->scratch("win_u", (size_t)rows * p * sizeof(double)));
        double* v = static_cast<double*>(g->scratch("win_v", (size_t)rows * p * sizeof(double)));
        for (int pass = 0; pass < 2; ++pass) {
            ContractArgs w{};
            w.out_f64 = pass == 0 ? u : v;
            w.ldo = p;
            w.cols_out = p;
            w.out_row0 = row0;
            run_phase2(g, row0, row0 + rows, pass == 0 ? c_dev : r_dev, pass == 0 ? ldc : ldr, 1, p, 6, EPI_WORK, w,
                       st, nullptr, nullptr);
        }
        double* G = static_cast<double*>(g->scratch("win_G", ((size_t)p * p + p) * sizeof(double)));
        double* csum = G + (size_t)p * p;
        launch_cov_gram(c_dev, ldc, g->n, p, G, csum, st);
        double* A = static_cast<double*>(g->scratch("win_A", (size_t)rows * 2 * p * sizeof(double)));
        double* B = static_cast<double*>(g->scratch("win_B", (size_t)rows * 2 * p * sizeof(double)));
        double* yb = static_cast<double*>(g->scratch("win_ybar", (size_t)rows * 2 * sizeof(double)));
        launch_win_cov_prep(u, v, p, G, csum, g->nobs, g->sum, g->sumsq, g->inv_sd, completed, row0, row0 + rows,
                            g->n, A, B, yb, yb + rows, st);
        a.scale = g->inv_sd;
        a.cov_a = A;
        a.cov_b = B;
        a.cov_k = 2 * p;
        a.ybar = yb;
        a.tsc = yb + rows;
    }
    const bool miss = g->impute_mode == GPE_IMPUTE_MEAN && g->n_missing > 0;
    ProfScope ps(GPE_PROF_GRAM, st);
    launch_gram_band(a, np, miss, st);
    ps.stop();
    // the host pair list and bounds must outlive the copies
    GPE_HIP(hipStreamSynchronize(st));
}

// Phase 1 entry: weights, probe streams (device MT19937-64 for Rademacher, host
// std::normal_distribution for real probes), then the traversal.  With
// v_start_dev the tile's first probe id is read on the device (Rademacher
// only), so a captured graph can advance it between replays.
void phase1_entry(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size, const double* annot_dev,
                  int32_t n_bins, const double* inv_dev, int32_t v_start, const int32_t* v_start_dev,
                  int32_t v_count, int32_t probe_tile, int32_t rand_dist, uint64_t root_seed, double* xz_dev,
                  int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, hipStream_t st) {
    require(probe_tile > 0, "probe_tile must be positive");
    require_prepared(g);
    require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "blk_end exceeds #SNPs in BIM");
    require(step_size > 0, "step_size must be positive");
    require(n_bins > 0 && v_count > 0, "annotation count and v_count must be positive");
    require(ldxz >= g->n, "Xz2d_chunk shape must be (N, B*v_count)");
    if (snp_end == snp_begin) return;
    (void)limb_cfg(limbs);
    CaptureScope cap(g, st);
    require(!cap.active || !prof().on, "kernel timing must be off during graph capture");
    require(!cap.active || rand_dist == GPE_RAND_RADEMACHER,
            "only Rademacher probes (generated on the device) can be captured in a graph");
    const int sub = std::min(probe_tile, v_count);
    const int64_t L = snp_end - snp_begin;
    const int v_pad = (int)round_up(v_count, 32);
    double* w = static_cast<double*>(g->scratch("p1_w", g->m_alloc * n_bins * sizeof(double)));
    hipLaunchKernelGGL(k_make_w, dim3((unsigned)ceil_div(L, 256)), dim3(256), 0, st, annot_dev, n_bins, inv_dev,
                       g->inv_sd, g->nobs, g->sum, g->sumsq, snp_begin, snp_end, w);
    GPE_HIP(hipGetLastError());
    std::vector<int64_t> starts;
    for (int64_t s = snp_begin; s < snp_end; s += step_size) starts.push_back(s);
    const int nb = (int)starts.size();
    int8_t* z_i8 = nullptr;
    double* z_f64 = nullptr;
    const int64_t ldz = v_pad;
    if (rand_dist == GPE_RAND_RADEMACHER) {
        z_i8 = static_cast<int8_t*>(g->scratch("p1_z8", g->m_alloc * ldz));
        int64_t* starts_dev = static_cast<int64_t*>(g->scratch("p1_starts", nb * sizeof(int64_t)));
        // the block starts stay on the device between calls with the same (begin, end, step)
        if (!(g->starts_key[0] == snp_begin && g->starts_key[1] == snp_end && g->starts_key[2] == step_size)) {
            require(!cap.active, "SNP-block starts not on the device: run the same call once outside graph capture first");
            GPE_HIP(hipMemcpyAsync(starts_dev, starts.data(), nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
            GPE_HIP(hipStreamSynchronize(st));
            g->starts_key[0] = snp_begin;
            g->starts_key[1] = snp_end;
            g->starts_key[2] = step_size;
        }
        launch_mt_rademacher(root_seed, starts_dev, nb, snp_end, v_start, v_count, z_i8, ldz, st, sub, v_start_dev);
    } else if (is_real_probe_dist(rand_dist)) {
        z_f64 = static_cast<double*>(g->scratch("p1_z64", g->m_alloc * ldz * sizeof(double)));
        // the previous call's copy out of the staging buffer must be done before it is rewritten
        GPE_HIP(hipStreamSynchronize(st));
        double* zp = g->stage((size_t)(L * ldz) * sizeof(double));
        // the host draws overlap the device: the SNP blocks are cut into parts, and each
        // part's copy and traversal are queued while the host draws the next part (later
        // parts add to what the first one wrote)
        const int parts = nb >= 16 ? 4 : 1;
        int b0 = 0;
        for (int p = 0; p < parts; ++p) {
            const int b1 = (int)((int64_t)nb * (p + 1) / parts);
            const int64_t ps = starts[b0], pe = b1 < nb ? starts[b1] : snp_end;
            double* zpp = zp + (ps - snp_begin) * ldz;
            for (int v0 = 0; v0 < v_count; v0 += sub) {
                const int vc = std::min(sub, v_count - v0);
                real_probes_into(rand_dist, root_seed, starts.data() + b0, b1 - b0, pe, v_start + v0, vc, zpp + v0,
                                 ldz);
            }
            if (v_pad > v_count)
                for (int64_t r = 0; r < pe - ps; ++r) std::fill(zpp + r * ldz + v_count, zpp + (r + 1) * ldz, 0.0);
            GPE_HIP(hipMemcpyAsync(z_f64 + ps * ldz, zpp, (size_t)((pe - ps) * ldz) * sizeof(double),
                                   hipMemcpyHostToDevice, st));
            run_phase1(g, ps, pe, w, n_bins, nullptr, z_f64, ldz, v_count, xz_dev, ldxz,
                       p == 0 ? bin_mode_dev : nullptr, limbs, st);
            b0 = b1;
        }
        return;
    } else {
        throw Error("rand_dist must be 'rademacher', 'gaussian' or 'spherical'");
    }
    run_phase1(g, snp_begin, snp_end, w, n_bins, z_i8, z_f64, ldz, v_count, xz_dev, ldxz, bin_mode_dev, limbs, st);
}

}  // namespace

extern "C" {

int gpe_phase1_xz(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size, const double* annot_dev,
                  int32_t n_bins, const double* inv_dev, int32_t v_start, int32_t v_count, int32_t rand_dist,
                  uint64_t root_seed, double* xz_dev, int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs,
                  void* stream) {
    return gpe_phase1_xz_tiled(g, snp_begin, snp_end, step_size, annot_dev, n_bins, inv_dev, v_start, v_count, v_count,
                               rand_dist, root_seed, xz_dev, ldxz, bin_mode_dev, limbs, stream);
}

int gpe_phase1_xz_tiled(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size, const double* annot_dev,
                        int32_t n_bins, const double* inv_dev, int32_t v_start, int32_t v_count, int32_t probe_tile,
                        int32_t rand_dist, uint64_t root_seed, double* xz_dev, int64_t ldxz,
                        const uint8_t* bin_mode_dev, int32_t limbs, void* stream) {
    return guarded([&] {
        phase1_entry(g, snp_begin, snp_end, step_size, annot_dev, n_bins, inv_dev, v_start, nullptr, v_count,
                     probe_tile, rand_dist, root_seed, xz_dev, ldxz, bin_mode_dev, limbs, S(stream));
    });
}

int gpe_phase1_xz_indirect(gpe_geno* g, int64_t snp_begin, int64_t snp_end, int64_t step_size,
                           const double* annot_dev, int32_t n_bins, const double* inv_dev,
                           const int32_t* v_start_dev, int32_t v_count, int32_t probe_tile, uint64_t root_seed,
                           double* xz_dev, int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream) {
    return guarded([&] {
        require(v_start_dev != nullptr, "v_start_dev is null");
        phase1_entry(g, snp_begin, snp_end, step_size, annot_dev, n_bins, inv_dev, 0, v_start_dev, v_count,
                     probe_tile, GPE_RAND_RADEMACHER, root_seed, xz_dev, ldxz, bin_mode_dev, limbs, S(stream));
    });
}

int gpe_add_i32(int32_t* value_dev, int32_t increment, void* stream) {
    return guarded([&] {
        require(value_dev != nullptr, "value_dev is null");
        launch_add_i32(value_dev, increment, S(stream));
    });
}

int gpe_xc_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* c_dev, int64_t ldc, int32_t n_cols,
               double* out_dev, int64_t ldo, int32_t accumulate, int32_t limbs, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "SNP range out of bounds");
        require(n_cols > 0 && ldc >= n_cols && ldo >= g->n, "bad leading dimension");
        if (snp_end == snp_begin) return;
        (void)limb_cfg(limbs);
        hipStream_t st = S(stream);
        const int64_t L = snp_end - snp_begin;
        const int v_pad = (int)round_up(n_cols, 32);
        double* w = static_cast<double*>(g->scratch("xc_w", g->m_alloc * sizeof(double)));
        hipLaunchKernelGGL(k_w_from_inv_sd, dim3((unsigned)ceil_div(L, 256)), dim3(256), 0, st, g->inv_sd, g->nobs,
                           g->sum, g->sumsq, snp_begin, snp_end, w);
        GPE_HIP(hipGetLastError());
        double* z = static_cast<double*>(g->scratch("xc_z", g->m_alloc * v_pad * sizeof(double)));
        copy_rows(c_dev, ldc, L, n_cols, z, v_pad, snp_begin, st);
        uint8_t* mode = static_cast<uint8_t*>(g->scratch("xc_mode", 1));
        launch_fill_u8(mode, 1, accumulate ? 2 : 1, st);
        run_phase1(g, snp_begin, snp_end, w, 1, nullptr, z, v_pad, n_cols, out_dev, ldo, mode, limbs, st);
    });
}

int gpe_xtv_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* v_dev, int64_t ldv, int32_t n_cols,
                double* out_dev, int64_t ldo, int32_t limbs, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "SNP range out of bounds");
        require(n_cols > 0 && ldv >= g->n && ldo >= n_cols, "bad leading dimension");
        if (snp_end == snp_begin) return;
        hipStream_t st = S(stream);
        ContractArgs a{};
        a.out_f64 = out_dev;
        a.ldo = ldo;
        a.cols_out = n_cols;
        a.out_row0 = snp_begin;
        run_phase2(g, snp_begin, snp_end, v_dev, ldv, 1, n_cols, limbs, EPI_WORK, a, st, nullptr, nullptr);
    });
}

int gpe_philox_rademacher(const uint64_t* keys_host, const int64_t* blk_starts_host, int32_t n_blocks, int64_t blk_end,
                          int32_t v_count, int8_t* z_dev, int64_t ldz, void* stream) {
    return guarded([&] {
        require(keys_host && blk_starts_host && z_dev && n_blocks > 0 && v_count > 0 && ldz >= v_count,
                "bad Philox probe request");
        hipStream_t st = S(stream);
        std::vector<int64_t> lens((size_t)n_blocks);
        int64_t max_len = 0;
        for (int b = 0; b < n_blocks; ++b) {
            const int64_t e = b + 1 < n_blocks ? blk_starts_host[b + 1] : blk_end;
            lens[(size_t)b] = e - blk_starts_host[b];
            require(lens[(size_t)b] > 0, "probe blocks must be non-empty and increasing");
            max_len = std::max(max_len, lens[(size_t)b]);
        }
        const size_t nkeys = (size_t)n_blocks * (size_t)v_count * 2;
        const size_t bytes = nkeys * sizeof(uint64_t) + 2 * (size_t)n_blocks * sizeof(int64_t);
        // per-device key / block tables (device + pinned host staging), kept for the
        // process and reused once the previous call's launch has read them
        struct Tables {
            uint8_t* dev = nullptr;
            uint8_t* host = nullptr;
            size_t bytes = 0;
            hipEvent_t done = nullptr;
        };
        static std::mutex mu;
        static std::map<int, Tables> tables;
        int device = 0;
        GPE_HIP(hipGetDevice(&device));
        std::lock_guard<std::mutex> lock(mu);
        Tables& t = tables[device];
        if (!t.done) GPE_HIP(hipEventCreateWithFlags(&t.done, hipEventDisableTiming));
        GPE_HIP(hipEventSynchronize(t.done));
        if (t.bytes < bytes) {
            if (t.dev) GPE_HIP(hipFree(t.dev));
            if (t.host) GPE_HIP(hipHostFree(t.host));
            t.dev = t.host = nullptr;
            t.bytes = 0;
            const size_t cap = std::max<size_t>(bytes, 1 << 16);
            GPE_HIP(hipMalloc(reinterpret_cast<void**>(&t.dev), cap));
            GPE_HIP(hipHostMalloc(reinterpret_cast<void**>(&t.host), cap, hipHostMallocDefault));
            t.bytes = cap;
        }
        std::memcpy(t.host, keys_host, nkeys * sizeof(uint64_t));
        std::memcpy(t.host + nkeys * sizeof(uint64_t), blk_starts_host, n_blocks * sizeof(int64_t));
        std::memcpy(t.host + nkeys * sizeof(uint64_t) + n_blocks * sizeof(int64_t), lens.data(),
                    n_blocks * sizeof(int64_t));
        GPE_HIP(hipMemcpyAsync(t.dev, t.host, bytes, hipMemcpyHostToDevice, st));
        const uint64_t* keys = reinterpret_cast<const uint64_t*>(t.dev);
        const int64_t* starts = reinterpret_cast<const int64_t*>(t.dev + nkeys * sizeof(uint64_t));
        const int64_t* lens_d = starts + n_blocks;
        launch_philox_rademacher(keys, starts, lens_d, n_blocks, v_count, max_len, blk_starts_host[0], z_dev, ldz, st);
        GPE_HIP(hipEventRecord(t.done, st));
    });
}

int gpe_phase1_xz_z8(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* annot_dev, int32_t n_bins,
                     const double* inv_dev, const int8_t* z_dev, int64_t ldz, int32_t v_count, double* xz_dev,
                     int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "blk_end exceeds #SNPs in BIM");
        require(n_bins > 0 && v_count > 0 && ldz >= v_count, "bad probe matrix");
        require(ldxz >= g->n, "Xz2d_chunk shape must be (N, B*v_count)");
        if (snp_end == snp_begin) return;
        (void)limb_cfg(limbs);
        hipStream_t st = S(stream);
        const int64_t L = snp_end - snp_begin;
        const int v_pad = (int)round_up(v_count, 32);
        // the per-call missing-call records hold each probe as a 2-bit sign and every
        // column of a bin shares one fixed-point scale: only -1, 0 and +1 are exact
        CaptureScope cap(g, st);
        if (!cap.active) {
            unsigned int* bad = static_cast<unsigned int*>(g->scratch("z8_check", sizeof(unsigned int)));
            launch_fill_u8(reinterpret_cast<uint8_t*>(bad), sizeof(unsigned int), 0, st);
            launch_z8_check(z_dev, ldz, L, v_count, bad, st);
            unsigned int bad_h = 0;
            GPE_HIP(hipMemcpyAsync(&bad_h, bad, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
            GPE_HIP(hipStreamSynchronize(st));
            require(bad_h == 0, "int8 probes must be -1, 0 or +1");
        }
        double* w = static_cast<double*>(g->scratch("p1_w", g->m_alloc * n_bins * sizeof(double)));
        hipLaunchKernelGGL(k_make_w, dim3((unsigned)ceil_div(L, 256)), dim3(256), 0, st, annot_dev, n_bins, inv_dev,
                           g->inv_sd, g->nobs, g->sum, g->sumsq, snp_begin, snp_end, w);
        GPE_HIP(hipGetLastError());
        int8_t* z = static_cast<int8_t*>(g->scratch("p1_z8", g->m_alloc * v_pad));
        GPE_HIP(hipMemcpy2DAsync(z + snp_begin * v_pad, v_pad, z_dev, ldz, v_count, L, hipMemcpyDeviceToDevice, st));
        run_phase1(g, snp_begin, snp_end, w, n_bins, z, nullptr, v_pad, v_count, xz_dev, ldxz, bin_mode_dev, limbs,
                   st);
    });
}

int gpe_phase1_xz_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* annot_dev, int32_t n_bins,
                      const double* inv_dev, const double* z_dev, int64_t ldz, int32_t v_count, double* xz_dev,
                      int64_t ldxz, const uint8_t* bin_mode_dev, int32_t limbs, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "blk_end exceeds #SNPs in BIM");
        require(n_bins > 0 && v_count > 0 && ldz >= v_count, "bad probe matrix");
        require(ldxz >= g->n, "Xz2d_chunk shape must be (N, B*v_count)");
        if (snp_end == snp_begin) return;
        (void)limb_cfg(limbs);
        hipStream_t st = S(stream);
        const int64_t L = snp_end - snp_begin;
        const int v_pad = (int)round_up(v_count, 32);
        double* w = static_cast<double*>(g->scratch("p1_w", g->m_alloc * n_bins * sizeof(double)));
        hipLaunchKernelGGL(k_make_w, dim3((unsigned)ceil_div(L, 256)), dim3(256), 0, st, annot_dev, n_bins, inv_dev,
                           g->inv_sd, g->nobs, g->sum, g->sumsq, snp_begin, snp_end, w);
        GPE_HIP(hipGetLastError());
        double* z = static_cast<double*>(g->scratch("p1_zc", g->m_alloc * v_pad * sizeof(double)));
        copy_rows(z_dev, ldz, L, v_count, z, v_pad, snp_begin, st);
        run_phase1(g, snp_begin, snp_end, w, n_bins, nullptr, z, v_pad, v_count, xz_dev, ldxz, bin_mode_dev, limbs,
                   st);
    });
}

int gpe_project_f64(double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* a_dev, int64_t lda,
                    const double* b_dev, int64_t ldb, int32_t p, int32_t center, void* stream) {
    return guarded([&] {
        require(n >= 0 && n_cols >= 0 && p >= 0, "projection sizes must be non-negative");
        require(n_cols <= 65535, "at most 65535 columns per projection call");
        if (n == 0 || n_cols == 0 || (p == 0 && !center)) return;
        require(x_dev != nullptr && ldx >= n, "projected matrix needs ldx >= n");
        if (p > 0) require(a_dev != nullptr && b_dev != nullptr && lda >= n && ldb >= n, "projection factor shapes");
        // per-device workspace, kept for the process (grown on demand)
        static std::mutex mu;
        static std::map<int, Buf> ws;
        int dev = 0;
        GPE_HIP(hipGetDevice(&dev));
        const size_t need = ((size_t)std::max(p, 1) * n_cols + 2 * (size_t)n_cols + (size_t)std::max(p, 1)) *
                            sizeof(double);
        std::lock_guard<std::mutex> lock(mu);
        Buf& b = ws[dev];
        if (b.bytes < need) {
            hipStream_t st0 = S(stream);
            GPE_HIP(hipStreamSynchronize(st0));
            if (b.p) GPE_HIP(hipFree(b.p));
            b.p = nullptr;
            b.bytes = 0;
            GPE_HIP(hipMalloc(&b.p, need));
            b.bytes = need;
        }
        double* T = static_cast<double*>(b.p);
        double* Ssum = T + (size_t)std::max(p, 1) * n_cols;
        double* mean = Ssum + n_cols;
        double* a = mean + n_cols;
        launch_project(x_dev, ldx, n, n_cols, a_dev, lda, b_dev, ldb, p, center != 0, T, Ssum, a, mean, S(stream));
    });
}

}  // extern "C"

namespace {

// Scratch of gpe_project_f64_noalloc: one fixed array per device, kept for the
// process.  Host code must never pass the symbol itself to a kernel (on the host
// it names a stand-in, not device memory): the device address is looked up once
// per device with hipGetSymbolAddress and only that pointer is used.
constexpr size_t kNoAllocScratchDoubles = (size_t)1 << 19;  // 4 MiB
__device__ __align__(256) double scratch[kNoAllocScratchDoubles];

double* noalloc_scratch_ptr() {
    static std::mutex mu;
    static std::map<int, double*> ptrs;
    int dev = 0;
    GPE_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    auto it = ptrs.find(dev);
    if (it != ptrs.end()) return it->second;
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || p == nullptr)
        throw Error(std::string("HIP error: hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    ptrs[dev] = static_cast<double*>(p);
    return static_cast<double*>(p);
}

}  // namespace

extern "C" {

int gpe_project_f64_noalloc(double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* a_dev,
                            int64_t lda, const double* b_dev, int64_t ldb, int32_t p, int32_t center,
                            void* stream) {
    return guarded([&] {
        require(n >= 0 && n_cols >= 0 && p >= 0, "projection sizes must be non-negative");
        require(n_cols <= 65535, "at most 65535 columns per projection call");
        const size_t need = (size_t)std::max(p, 1) * (size_t)n_cols + 2 * (size_t)n_cols + (size_t)std::max(p, 1);
        require(need <= kNoAllocScratchDoubles,
                "projection needs " + std::to_string(need) + " scratch doubles; the allocation-free entry holds " +
                    std::to_string(kNoAllocScratchDoubles) + " ((max(p,1)+2) * n_cols + max(p,1) <= 524288)");
        if (n == 0 || n_cols == 0 || (p == 0 && !center)) return;
        require(x_dev != nullptr && ldx >= n, "projected matrix needs ldx >= n");
        if (p > 0) require(a_dev != nullptr && b_dev != nullptr && lda >= n && ldb >= n, "projection factor shapes");
        // every scratch value is written by this call's kernels before it is read
        // (coefficients, column sums and means), so nothing is cleared
        double* T = noalloc_scratch_ptr();
        double* Ssum = T + (size_t)std::max(p, 1) * n_cols;
        double* mean = Ssum + n_cols;
        double* a = mean + n_cols;
        launch_project(x_dev, ldx, n, n_cols, a_dev, lda, b_dev, ldb, p, center != 0, T, Ssum, a, mean, S(stream));
    });
}

int gpe_col_sums_f64(const double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, double* out_dev, void* stream) {
    return guarded([&] {
        require(n >= 0 && n_cols >= 0, "column-sum sizes must be non-negative");
        require(n_cols <= 2147483647 && (n_cols == 0 || (x_dev != nullptr && out_dev != nullptr && ldx >= n)),
                "column sums need ldx >= n");
        if (n_cols == 0) return;
        launch_col_sums(x_dev, ldx, n, n_cols, out_dev, S(stream));
    });
}

int gpe_basis_coef_f64(const double* x_dev, int64_t ldx, int64_t n, int32_t n_cols, const double* b_dev, int64_t ldb,
                       int32_t p, double* t_dev, int64_t ldt, double* sumsq_dev, void* stream) {
    return guarded([&] {
        require(n >= 0 && n_cols >= 0 && p >= 0, "basis-coefficient sizes must be non-negative");
        if (n == 0 || n_cols == 0) return;
        require(x_dev != nullptr && ldx >= n, "basis coefficients need ldx >= n");
        if (p > 0) require(b_dev != nullptr && ldb >= n && t_dev != nullptr && ldt >= n_cols, "basis factor shapes");
        launch_basis_coef(x_dev, ldx, n, n_cols, b_dev, ldb, p, t_dev, ldt, sumsq_dev, S(stream));
    });
}

int gpe_sq_rowsums_f64(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* w_dev, int64_t ldw,
                       int32_t n_bins, const double* c_dev, int64_t ldc, const double* rx_dev, int64_t ldrx, int32_t p,
                       double* out_dev, int64_t ldo, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "SNP range out of bounds");
        const int64_t L = snp_end - snp_begin;
        require(n_bins >= 0 && (n_bins == 0 || (w_dev != nullptr && out_dev != nullptr && ldw >= L && ldo >= g->n)),
                "bad squared-row-sum shape");
        require(p >= 0 && p <= 64, "squared row sums take at most 64 covariates");
        if (p > 0)
            require(c_dev != nullptr && rx_dev != nullptr && ldc >= g->n && ldrx >= p, "C/R shape mismatch");
        if (n_bins == 0 || L == 0) return;
        const int64_t ny = ceil_div(L, sq_rowsums_chunk(g->n, L));
        double* part = static_cast<double*>(g->scratch("sq_part", (size_t)ny * 8 * g->n * sizeof(double)));
        launch_sq_rowsums(g->smp_work, g->smp_stride, g->n, snp_begin, snp_end, g->mean, g->inv_sd, w_dev, ldw, n_bins,
                          c_dev, ldc, rx_dev, ldrx, p, part, out_dev, ldo, S(stream));
    });
}

int gpe_weighted_sumsq(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* w_dev, int32_t n_weights,
                       int64_t ldw, double* out_dev, int64_t ldo, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "SNP range out of bounds");
        require(n_weights >= 0 && ldw >= g->n && ldo >= snp_end - snp_begin, "bad weighted-sum-of-squares shape");
        if (n_weights == 0 || snp_end == snp_begin) return;
        // int32 limb sums over N samples of squared counts (<= 4) times balanced digits (|d| <= 128)
        require(g->n <= 4000000, "weighted sums of squares support at most 4,000,000 samples");
        hipStream_t st = S(stream);
        // x_jn^2 = inv_sd_j^2 (g_jn - mu_j)^2 on observed calls (0 where missing, mean mode), and g^2 = g + 2[g=2],
        // so sum_n w_n x_jn^2 = inv_sd_j^2 (mu_j^2 (W - ME) - 2 mu_j GE + G2E) with exact int64 plane sums of the
        // quantized weights: GE (counts), ME (missing flags), G2E (squared counts), W (all samples)
        const LimbCfg lc = limb_cfg(6);
        const int64_t L = snp_end - snp_begin;
        for (int e0 = 0; e0 < n_weights; e0 += 1024) {
            const int nw = std::min(1024, n_weights - e0);
            const int v_pad = (int)round_up(nw, 32);
            const int groups = v_pad / 32;
            const int64_t coef_ctiles = (int64_t)groups * lc.nl * lc.ctc;
            const int64_t kcount = ceil_div(g->n, kKChunk);
            uint64_t* colmax = static_cast<uint64_t*>(g->scratch("colmax", v_pad * sizeof(uint64_t)));
            int32_t* shift = static_cast<int32_t*>(g->scratch("shift", v_pad * sizeof(int32_t)));
            double* col_scale = static_cast<double*>(g->scratch("col_scale", v_pad * sizeof(double)));
            int64_t* col_sum = static_cast<int64_t*>(g->scratch("col_sum", v_pad * sizeof(int64_t)));
            launch_fill_u64(colmax, v_pad, 0, st);
            launch_fill_u64(reinterpret_cast<uint64_t*>(col_sum), v_pad, 0, st);
            const double* w = w_dev + (int64_t)e0 * ldw;
            launch_colmax_f64(w, ldw, g->n, nw, nw, v_pad, colmax, st);
            const int prec = std::min(8 * lc.nl - 2, 62 - bitlen((uint64_t)(2 * g->n)));
            launch_shift_from_colmax(colmax, v_pad, prec, shift, col_scale, st);
            int8_t* coef = static_cast<int8_t*>(g->scratch("coef", kcount * coef_ctiles * 4096));
            if (v_pad != nw) launch_fill_u8(reinterpret_cast<uint8_t*>(coef), kcount * coef_ctiles * 4096, 0, st);
            launch_quant_p2(w, ldw, g->n, 1, nw, v_pad, shift, lc.nl, lc.ctc, coef_ctiles, coef, col_sum, st);
            int64_t* ge = static_cast<int64_t*>(g->scratch("wss_ge", (size_t)L * v_pad * sizeof(int64_t)));
            int64_t* me = static_cast<int64_t*>(g->scratch("wss_me", (size_t)L * v_pad * sizeof(int64_t)));
            int64_t* g2 = static_cast<int64_t*>(g->scratch("wss_g2", (size_t)L * v_pad * sizeof(int64_t)));
            ContractArgs a{};
            a.packed = g->snp_tiled ? g->snp_tiled : g->snp_work;
            a.tiled = g->snp_tiled ? 1 : 0;
            a.stride = g->snp_stride;
            a.r0 = snp_begin;
            a.r1 = snp_end;
            a.coef = coef;
            a.coef_ctiles = coef_ctiles;
            a.kc0 = 0;
            a.kc1 = kcount;
            a.kc_per_split = kcount;
            a.ctiles_used = v_pad / 32;
            a.ldo = v_pad;
            a.cols_out = v_pad;
            a.out_row0 = snp_begin;
            const bool miss = g->impute_mode == GPE_IMPUTE_MEAN && g->n_missing > 0;
            a.out_i64 = ge;
            a.out2_i64 = me;
            launch_contract(lc.nl, lc.ctc, miss, miss ? EPI_I64_2P : EPI_I64_ROWMAJOR, a, groups, 1, st);
            a.out_i64 = g2;
            a.out2_i64 = nullptr;
            launch_contract(lc.nl, lc.ctc, false, EPI_I64_SQ, a, groups, 1, st);
            launch_wss_finalize(ge, miss ? me : nullptr, g2, v_pad, snp_begin, snp_end, nw, g->mean, g->inv_sd,
                                col_scale, col_sum, out_dev + (int64_t)e0 * ldo, ldo, st);
        }
    });
}

int gpe_windowed_ld(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                    int64_t window_snps, const double* annot_dev, int32_t n_bins, double* out_dev, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_end >= snp_begin, "chr_end must be >= chr_start");
        require(snp_end - snp_begin > 0, "Chromosome slice is empty");
        const int64_t bounds[2] = {snp_begin, snp_end};
        windowed_ld_impl(g, 1, bounds, bp_host, window_bp, window_snps, annot_dev, n_bins, out_dev, S(stream));
    });
}

int gpe_windowed_ld_side(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                         int64_t window_snps, int32_t side, const double* annot_dev, int32_t n_bins, double* out_dev,
                         void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(snp_end - snp_begin > 0, "Chromosome slice is empty");
        const int64_t b[2] = {snp_begin, snp_end};
        windowed_ld_impl(g, 1, b, bp_host, window_bp, window_snps, annot_dev, n_bins, out_dev, S(stream), nullptr, 0,
                         nullptr, 0, 0, side);
    });
}

int gpe_windowed_ld_cov(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const int64_t* bp_host, double window_bp,
                        int64_t window_snps, const double* annot_dev, int32_t n_bins, const double* c_dev,
                        int64_t ldc, const double* r_dev, int64_t ldr, int32_t p, double* out_dev, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(p > 0, "C/R shape mismatch in compute_windowed_ld_chr");
        const int64_t b[2] = {snp_begin, snp_end};
        windowed_ld_impl(g, 1, b, bp_host, window_bp, window_snps, annot_dev, n_bins, out_dev, S(stream), c_dev, ldc,
                         r_dev, ldr, p);
    });
}

int gpe_windowed_ld_multi(gpe_geno* g, int32_t n_chr, const int64_t* chr_bounds_host, const int64_t* bp_host,
                          double window_bp, int64_t window_snps, const double* annot_dev, int32_t n_bins,
                          double* out_dev, void* stream) {
    return guarded([&] {
        require_prepared(g);
        require(n_chr > 0 && chr_bounds_host != nullptr, "at least one chromosome is required");
        windowed_ld_impl(g, n_chr, chr_bounds_host, bp_host, window_bp, window_snps, annot_dev, n_bins, out_dev,
                         S(stream));
    });
}

}  // extern "C"

namespace {

void phase2_entry(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* inv_left_dev, int32_t tile_nvecs,
                  const double* xz_dev, int64_t ldxz, int32_t n_bins, int32_t n_denom, double* meansq_dev,
                  int64_t ld_meansq, double* mc_aggregate_sumsq_dev, double* mc_per_snp_sumsq_dev,
                  double extreme_warn, int32_t drop_extreme, int32_t limbs, const double* sum_xz_dev,
                  unsigned long long* guard_counts_host, void* stream) {
    {
        require_prepared(g);
        require(snp_begin >= 0 && snp_end <= g->m && snp_begin <= snp_end, "blk_end exceeds #SNPs in BIM");
        require(tile_nvecs > 0 && n_bins > 0, "Xz2d col count must be a multiple of tile_nvecs");
        require(ldxz >= g->n, "Xz2d shape mismatch in phase2_accum_XtXz_bed");
        require(ld_meansq >= n_bins, "meansq_accum shape mismatch in phase2_accum_XtXz_bed");
        hipStream_t st = S(stream);
        CaptureScope cap(g, st);
        require(!cap.active || (guard_counts_host == nullptr && !prof().on),
                "under graph capture the guard counts cannot be returned and kernel timing must be off");
        unsigned long long* gc = static_cast<unsigned long long*>(g->scratch("guards", 2 * sizeof(unsigned long long)));
        launch_fill_u64(reinterpret_cast<uint64_t*>(gc), 2, 0, st);
        if (snp_end > snp_begin) {
            (void)limb_cfg(limbs);
            const int v_pad = (int)round_up(tile_nvecs, 32);
            const int ctiles_used = n_bins * v_pad / 32;
            const bool mc = mc_aggregate_sumsq_dev || mc_per_snp_sumsq_dev;
            ContractArgs a{};
            a.s2 = static_cast<double*>(g->scratch("p2_s2", g->m_alloc * ctiles_used * sizeof(double)));
            a.s4 = mc ? static_cast<double*>(g->scratch("p2_s4", g->m_alloc * ctiles_used * sizeof(double))) : nullptr;
            a.ld_s = ctiles_used;
            run_phase2(g, snp_begin, snp_end, xz_dev, ldxz, n_bins, tile_nvecs, limbs, EPI_P2, a, st, nullptr,
                       nullptr, sum_xz_dev);
            double denom = (double)n_denom - 1.0;
            if (denom <= 0.0) denom = 1.0;
            const double inv_denom2 = 1.0 / (denom * denom);
            launch_p2_finalize(a.s2, a.s4, ctiles_used, snp_begin, snp_end, n_bins, v_pad / 32, tile_nvecs,
                               inv_left_dev, inv_denom2, extreme_warn > 0.0 ? extreme_warn : INFINITY, drop_extreme,
                               meansq_dev, ld_meansq, mc_aggregate_sumsq_dev, mc_per_snp_sumsq_dev, gc, st);
        }
        if (guard_counts_host) {
            GPE_HIP(hipMemcpyAsync(guard_counts_host, gc, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
            GPE_HIP(hipStreamSynchronize(st));
        }
    }
}

}  // namespace

extern "C" {

int gpe_phase2_accum(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* inv_left_dev,
                     int32_t tile_nvecs, const double* xz_dev, int64_t ldxz, int32_t n_bins, int32_t n_denom,
                     double* meansq_dev, int64_t ld_meansq, double* mc_aggregate_sumsq_dev,
                     double* mc_per_snp_sumsq_dev, double extreme_warn, int32_t drop_extreme, int32_t limbs,
                     unsigned long long* guard_counts_host, void* stream) {
    return guarded([&] {
        phase2_entry(g, snp_begin, snp_end, inv_left_dev, tile_nvecs, xz_dev, ldxz, n_bins, n_denom, meansq_dev,
                     ld_meansq, mc_aggregate_sumsq_dev, mc_per_snp_sumsq_dev, extreme_warn, drop_extreme, limbs,
                     nullptr, guard_counts_host, stream);
    });
}

int gpe_phase2_accum_colsums(gpe_geno* g, int64_t snp_begin, int64_t snp_end, const double* inv_left_dev,
                             int32_t tile_nvecs, const double* xz_dev, int64_t ldxz, int32_t n_bins,
                             const double* sum_xz_dev, int32_t n_denom, double* meansq_dev, int64_t ld_meansq,
                             double* mc_aggregate_sumsq_dev, double* mc_per_snp_sumsq_dev, double extreme_warn,
                             int32_t drop_extreme, int32_t limbs, unsigned long long* guard_counts_host,
                             void* stream) {
    return guarded([&] {
        require(sum_xz_dev != nullptr, "sum_Xz shape mismatch in phase2_accum_XtXz_bed_mailman_rm");
        phase2_entry(g, snp_begin, snp_end, inv_left_dev, tile_nvecs, xz_dev, ldxz, n_bins, n_denom, meansq_dev,
                     ld_meansq, mc_aggregate_sumsq_dev, mc_per_snp_sumsq_dev, extreme_warn, drop_extreme, limbs,
                     sum_xz_dev, guard_counts_host, stream);
    });
}

}  // extern "C"
