// ORACLE — test infrastructure only (never linked into the product).
//
// C-ABI harness around the reference's own BED decoder, compiled from the
// reference sources where they lie (src/native/common/genotype.cpp, which is
// self-contained: libc + the C++ standard library).  Used only to pin the
// oracle's restatement of decode / standardize / HWE imputation
// (src/native/common/genotype.cpp:474-952) and to write golden vectors.
//
// It also exposes libstdc++'s std::mt19937_64 / std::normal_distribution with
// the reference's seed derivation (src/native/gwldcore.cpp:445-458,
// :1679-1716), the generator library the reference itself calls, to pin the
// oracle's plain-C restatement of those streams.
#include "genotype.hpp"

#include <cstdint>
#include <cstring>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

static thread_local std::string g_err;

extern "C" const char* ref_last_error() { return g_err.c_str(); }

extern "C" int ref_read_block_standardized_double(const char* bed_path, const char* fam_path,
                                                  int blk_start, int blk_end, const int* rows,
                                                  int nrows, int ddof, int impute_mode,
                                                  uint64_t impute_seed, double* out) {
    try {
        std::vector<int> r(rows, rows + nrows);
        std::vector<double> geno;
        int N = 0, L = 0;
        read_block_standardized_double(bed_path, fam_path, blk_start, blk_end, r, ddof,
                                       impute_mode == 1 ? ImputeMode::Hwe : ImputeMode::Mean,
                                       impute_seed, geno, N, L);
        std::memcpy(out, geno.data(), sizeof(double) * (size_t)N * (size_t)L);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

// The reference's in-place standardized decode (the call its phase 1 and phase 2
// make per block, src/native/gwldcore.cpp:1582-1861, :2329-2600) in float
// (is_float != 0) or double into out (N x L column-major, capacity elems).
extern "C" int ref_read_block_standardized_into(const char* bed_path, const char* fam_path, int blk_start,
                                                int blk_end, const int* rows, int nrows, int ddof, int impute_mode,
                                                uint64_t impute_seed, void* out, size_t elems, int is_float) {
    try {
        std::vector<int> r(rows, rows + nrows);
        int N = 0, L = 0;
        const ImputeMode mode = impute_mode == 1 ? ImputeMode::Hwe : ImputeMode::Mean;
        if (is_float)
            read_block_standardized_into<float>(bed_path, fam_path, blk_start, blk_end, r, ddof, mode, impute_seed,
                                                static_cast<float*>(out), elems, N, L);
        else
            read_block_standardized_into<double>(bed_path, fam_path, blk_start, blk_end, r, ddof, mode, impute_seed,
                                                 static_cast<double*>(out), elems, N, L);
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

extern "C" int ref_read_block_mailman_hwe(const char* bed_path, const char* fam_path,
                                          int blk_start, int blk_end, const int* rows, int nrows,
                                          int ddof, uint64_t impute_seed, double* mean_out,
                                          double* inv_std_out, int* observed_out) {
    try {
        std::vector<int> r(rows, rows + nrows);
        MailmanPackedBlock pack;
        read_block_mailman_hwe(bed_path, fam_path, blk_start, blk_end, r, ddof, impute_seed, pack);
        for (int i = 0; i < pack.L; ++i) {
            mean_out[i] = pack.mean[(size_t)i];
            inv_std_out[i] = pack.inv_std[(size_t)i];
            observed_out[i] = pack.observed.empty() ? -1 : pack.observed[(size_t)i];
        }
        return 0;
    } catch (const std::exception& e) {
        g_err = e.what();
        return 1;
    }
}

static inline uint64_t h_mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

static inline uint64_t h_make_seed(uint64_t root, int block, int v0) {
    uint64_t s = 0x1234abcdULL;
    s ^= h_mix64(root);
    s ^= h_mix64(static_cast<uint64_t>(block) + 0x9e37ULL);
    s ^= h_mix64(static_cast<uint64_t>(v0) + 0x85ebULL);
    return s;
}

// The reference's probe fill (src/native/gwldcore.cpp:1691-1716) with the
// standard library it is built against, for T = double and T = float.
template <class T>
static void h_probe_fill(uint64_t root, int blk_start, int v_start, int L, int V, bool rademacher, bool spherical,
                         double* Z) {
    std::mt19937_64 rng(h_make_seed(root, blk_start, v_start));
    std::normal_distribution<T> gN(0, (T)1);
    std::vector<T> zc((size_t)L);
    for (int c = 0; c < V; ++c) {
        long double ss = 0.0L;
        for (int r = 0; r < L; ++r) {
            T z = rademacher ? ((rng() & 1) ? T(+1) : T(-1)) : gN(rng);
            zc[(size_t)r] = z;
            if (spherical) ss += (long double)z * (long double)z;
        }
        if (spherical) {
            const T zscale = ss > 0.0L ? (T)std::sqrt((long double)L / ss) : T(1);
            for (int r = 0; r < L; ++r) zc[(size_t)r] *= zscale;
        }
        for (int r = 0; r < L; ++r) Z[(size_t)c * (size_t)L + (size_t)r] = (double)zc[(size_t)r];
    }
}

// dist: 0 rademacher, 1 gaussian, 2 spherical (double); 3 gaussian, 4 spherical (float)
extern "C" void ref_std_probe_block(uint64_t root, int blk_start, int v_start, int L, int V,
                                    int dist, double* Z) {
    if (dist >= 3)
        h_probe_fill<float>(root, blk_start, v_start, L, V, false, dist == 4, Z);
    else
        h_probe_fill<double>(root, blk_start, v_start, L, V, dist == 0, dist == 2, Z);
}
