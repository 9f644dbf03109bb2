/*
 * ORACLE — test infrastructure only.  Nothing in the product links this file;
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it
 * (through oracle/oracle.py), and only as the checker.
 *
 * Plain-C restatement of the probe streams of the genome-wide LD-score engine:
 *
 *   - MT19937-64, written from the published algorithm (Matsumoto & Nishimura,
 *     "mt19937-64", 2004) with the parameters the C++ standard fixes for
 *     std::mt19937_64 ([rand.predef]).  The reference draws every genome-wide
 *     probe from std::mt19937_64 (src/native/gwldcore.cpp:1679).
 *     Pinned by the standard's known answer: the 10000th output of a
 *     default-seeded engine is 9981545732273789042.
 *   - mix64 / make_seed: src/native/gwldcore.cpp:445-458.
 *   - Rademacher fill, column by column, (rng() & 1) ? +1 : -1:
 *     src/native/gwldcore.cpp:1691-1716.
 *   - Gaussian fill: libstdc++'s std::normal_distribution (Marsaglia polar
 *     method fed by std::generate_canonical), the documented algorithm of the
 *     GNU implementation the reference is built with, for T = double and
 *     T = float (the reference draws in the precision of the caller's arrays).
 *   - Spherical fill: the Gaussian fill with each probe column rescaled to
 *     norm sqrt(L), the sum of squares in long double:
 *     src/native/gwldcore.cpp:1702-1715.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#define MT_NN 312
#define MT_MM 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UM 0xFFFFFFFF80000000ULL
#define MT_LM 0x7FFFFFFFULL

typedef struct {
    uint64_t mt[MT_NN];
    int mti;
} or_mt64;

void or_mt64_seed(or_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < MT_NN; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->mti = MT_NN;
}

static void or_mt64_twist(or_mt64* s) {
    uint64_t* mt = s->mt;
    int i;
    uint64_t x;
    for (i = 0; i < MT_NN - MT_MM; ++i) {
        x = (mt[i] & MT_UM) | (mt[i + 1] & MT_LM);
        mt[i] = mt[i + MT_MM] ^ (x >> 1) ^ ((x & 1ULL) ? MT_MATRIX_A : 0ULL);
    }
    for (; i < MT_NN - 1; ++i) {
        x = (mt[i] & MT_UM) | (mt[i + 1] & MT_LM);
        mt[i] = mt[i + (MT_MM - MT_NN)] ^ (x >> 1) ^ ((x & 1ULL) ? MT_MATRIX_A : 0ULL);
    }
    x = (mt[MT_NN - 1] & MT_UM) | (mt[0] & MT_LM);
    mt[MT_NN - 1] = mt[MT_MM - 1] ^ (x >> 1) ^ ((x & 1ULL) ? MT_MATRIX_A : 0ULL);
    s->mti = 0;
}

uint64_t or_mt64_next(or_mt64* s) {
    if (s->mti >= MT_NN) or_mt64_twist(s);
    uint64_t x = s->mt[s->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

/* n-th output (1-based) of an engine seeded with `seed`. */
uint64_t or_mt64_nth(uint64_t seed, int64_t n) {
    or_mt64 s;
    or_mt64_seed(&s, seed);
    uint64_t v = 0;
    for (int64_t k = 0; k < n; ++k) v = or_mt64_next(&s);
    return v;
}

/* src/native/gwldcore.cpp:445-451 */
uint64_t or_mix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ULL;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    x = x ^ (x >> 31);
    return x;
}

/* src/native/gwldcore.cpp:452-458 */
uint64_t or_make_seed(uint64_t root, int block, int v0) {
    uint64_t s = 0x1234abcdULL;
    s ^= or_mix64(root);
    s ^= or_mix64((uint64_t)(int64_t)block + 0x9e37ULL);
    s ^= or_mix64((uint64_t)(int64_t)v0 + 0x85ebULL);
    return s;
}

/* Rademacher probes, column-major L x V: src/native/gwldcore.cpp:1691-1716. */
void or_rademacher_block(uint64_t root, int blk_start, int v_start, int L, int V, double* Z) {
    or_mt64 s;
    or_mt64_seed(&s, or_make_seed(root, blk_start, v_start));
    for (int c = 0; c < V; ++c)
        for (int r = 0; r < L; ++r)
            Z[(size_t)c * (size_t)L + (size_t)r] = (or_mt64_next(&s) & 1ULL) ? 1.0 : -1.0;
}

/* std::generate_canonical<double, 53>(mt19937_64): one 64-bit draw scaled by
 * 2^-64; a result that rounds to 1.0 is replaced by the largest double below 1. */
static double or_canonical_double(or_mt64* s) {
    const double r = 18446744073709551616.0; /* 2^64 */
    double ret = (double)or_mt64_next(s) / r;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

/* Marsaglia polar method as libstdc++ implements normal_distribution<double>:
 * one accepted pair yields two variates; the second is cached for the next
 * call on the same distribution object. */
typedef struct {
    int saved_available;
    double saved;
} or_normal;

static double or_normal_next(or_normal* nd, or_mt64* s) {
    if (nd->saved_available) {
        nd->saved_available = 0;
        return nd->saved;
    }
    double x, y, r2;
    do {
        x = 2.0 * or_canonical_double(s) - 1.0;
        y = 2.0 * or_canonical_double(s) - 1.0;
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const double mult = sqrt(-2.0 * log(r2) / r2);
    nd->saved = x * mult;
    nd->saved_available = 1;
    return y * mult;
}

/* Gaussian probes (double), column-major L x V, one normal_distribution
 * object shared by the whole fill: src/native/gwldcore.cpp:1680, :1705. */
void or_gaussian_block(uint64_t root, int blk_start, int v_start, int L, int V, double* Z) {
    or_mt64 s;
    or_normal nd = {0, 0.0};
    or_mt64_seed(&s, or_make_seed(root, blk_start, v_start));
    for (int c = 0; c < V; ++c)
        for (int r = 0; r < L; ++r)
            Z[(size_t)c * (size_t)L + (size_t)r] = or_normal_next(&nd, &s);
}

/* std::generate_canonical<float, 24>(mt19937_64): one 64-bit draw converted to
 * float (round to nearest) and divided by 2^64 in float. */
static float or_canonical_float(or_mt64* s) {
    const float r = 18446744073709551616.0f; /* 2^64 */
    float ret = (float)or_mt64_next(s) / r;
    if (ret >= 1.0f) ret = nextafterf(1.0f, 0.0f);
    return ret;
}

typedef struct {
    int saved_available;
    float saved;
} or_normalf;

/* normal_distribution<float>: 2 * U is a float product, the "- 1.0" is a double
 * subtraction rounded back to float, the rest is float arithmetic (logf, sqrtf). */
static float or_normalf_next(or_normalf* nd, or_mt64* s) {
    if (nd->saved_available) {
        nd->saved_available = 0;
        return nd->saved;
    }
    float x, y, r2;
    do {
        x = (float)((double)(2.0f * or_canonical_float(s)) - 1.0);
        y = (float)((double)(2.0f * or_canonical_float(s)) - 1.0);
        r2 = x * x + y * y;
    } while (r2 > 1.0 || r2 == 0.0);
    const float mult = sqrtf(-2.0f * logf(r2) / r2);
    nd->saved = x * mult;
    nd->saved_available = 1;
    return y * mult;
}

/* dist: 1 gaussian, 2 spherical (double); 3 gaussian, 4 spherical (float, stored
 * as double).  Column-major L x V; one distribution object for the whole fill. */
void or_real_probe_block(uint64_t root, int blk_start, int v_start, int L, int V, int dist, double* Z) {
    or_mt64 s;
    or_normal nd = {0, 0.0};
    or_normalf ndf = {0, 0.0f};
    or_mt64_seed(&s, or_make_seed(root, blk_start, v_start));
    const int f32 = dist >= 3;
    const int spherical = dist == 2 || dist == 4;
    for (int c = 0; c < V; ++c) {
        double* zc = Z + (size_t)c * (size_t)L;
        long double ss = 0.0L;
        if (f32) {
            for (int r = 0; r < L; ++r) {
                const float z = or_normalf_next(&ndf, &s);
                zc[r] = (double)z;
                if (spherical) ss += (long double)z * (long double)z;
            }
            if (spherical) {
                const float sc = ss > 0.0L ? (float)sqrtl((long double)L / ss) : 1.0f;
                for (int r = 0; r < L; ++r) zc[r] = (double)((float)zc[r] * sc);
            }
        } else {
            for (int r = 0; r < L; ++r) {
                const double z = or_normal_next(&nd, &s);
                zc[r] = z;
                if (spherical) ss += (long double)z * (long double)z;
            }
            if (spherical) {
                const double sc = ss > 0.0L ? (double)sqrtl((long double)L / ss) : 1.0;
                for (int r = 0; r < L; ++r) zc[r] = zc[r] * sc;
            }
        }
    }
}

/* Raw engine outputs for a (root, block, v0) stream, for stream tests. */
void or_stream_raw(uint64_t root, int blk_start, int v_start, int64_t n, uint64_t* out) {
    or_mt64 s;
    or_mt64_seed(&s, or_make_seed(root, blk_start, v_start));
    for (int64_t k = 0; k < n; ++k) out[k] = or_mt64_next(&s);
}
