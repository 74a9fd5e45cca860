/* orc_rng.c -- oracle random numbers.  TEST INFRASTRUCTURE ONLY.
 *
 * MT19937-64 (Nishimura & Matsumoto, ACM TOMACS 10 (2000) 348) with the
 * reference's seeding and transforms: Random.cpp:67-73 (init_genrand64),
 * 112-145 (genrand64_int64), 153-165 (real1/2/3), 178-204 (polar Gauss with
 * the cached second value).
 * Pinned by: std::mt19937_64 outputs (same recurrence and seeding) and the
 * C++ standard's KAT (10000th output of seed 5489); the Gauss and real
 * transforms are closed forms of the recorded raw words.
 *
 * Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and the project's counter
 * map for colour-charge normals.  Pinned by the Random123 known-answer
 * vectors.  The reference draws rho from the MT stream (Init.cpp:1697-1705);
 * the counter map is this project's parallel-friendly replacement, and the MT
 * stream mode of orc_setv keeps the reference order available.
 */
#include <math.h>
#include <string.h>

#include "ipg_oracle.h"
#include "orc_internal.h"

#define MT_N 312
#define MT_M 156
#define MT_MATRIX_A 0xB5026F5AA96619E9ULL
#define MT_UPPER 0xFFFFFFFF80000000ULL
#define MT_LOWER 0x000000007FFFFFFFULL

void orc_mt_init(orc_mt *r, uint64_t seed) {
    r->mt[0] = seed;
    for (int i = 1; i < MT_N; i++) {
        uint64_t prev = r->mt[i - 1];
        r->mt[i] = 6364136223846793005ULL * (prev ^ (prev >> 62)) + (uint64_t)i;
    }
    r->mti = MT_N;
    r->iset = 0;
    r->gset = 0.0;
}

static void mt_refill(orc_mt *r) {
    for (int i = 0; i < MT_N; i++) {
        uint64_t y = (r->mt[i] & MT_UPPER) | (r->mt[(i + 1) % MT_N] & MT_LOWER);
        uint64_t v = r->mt[(i + MT_M) % MT_N] ^ (y >> 1);
        if (y & 1ULL) v ^= MT_MATRIX_A;
        r->mt[i] = v;
    }
    r->mti = 0;
}

uint64_t orc_mt_u64(orc_mt *r) {
    if (r->mti >= MT_N) mt_refill(r);
    uint64_t x = r->mt[r->mti++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= (x >> 43);
    return x;
}

double orc_mt_real1(orc_mt *r) {
    return (double)(orc_mt_u64(r) >> 11) * (1.0 / 9007199254740991.0);
}
double orc_mt_real2(orc_mt *r) {
    return (double)(orc_mt_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}
double orc_mt_real3(orc_mt *r) {
    return ((double)(orc_mt_u64(r) >> 12) + 0.5) * (1.0 / 4503599627370496.0);
}

/* Marsaglia polar method; the second deviate is cached (Random.cpp:178-204) */
double orc_mt_gauss(orc_mt *r) {
    if (r->iset) {
        r->iset = 0;
        return r->gset;
    }
    double v1, v2, rsq;
    do {
        v1 = 2.0 * orc_mt_real3(r) - 1.0;
        v2 = 2.0 * orc_mt_real3(r) - 1.0;
        rsq = v1 * v1 + v2 * v2;
    } while (rsq > 1.0 || rsq == 0.0);
    double fac = sqrt(-2.0 * log(rsq) / rsq);
    r->gset = v1 * fac;
    r->iset = 1;
    return v2 * fac;
}

void orc_mt_fill(orc_mt *r, int kind, int64_t n, double *out) {
    for (int64_t i = 0; i < n; i++) {
        switch (kind) {
            case 0: {
                uint64_t u = orc_mt_u64(r);
                memcpy(&out[i], &u, 8);
                break;
            }
            case 1: out[i] = orc_mt_real1(r); break;
            case 2: out[i] = orc_mt_real2(r); break;
            case 3: out[i] = orc_mt_real3(r); break;
            default: out[i] = orc_mt_gauss(r); break;
        }
    }
}

/* ---- Philox4x32-10 ------------------------------------------------------ */
#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; round++) {
        if (round > 0) {
            k0 += PHILOX_W0;
            k1 += PHILOX_W1;
        }
        uint64_t p0 = (uint64_t)PHILOX_M0 * c0;
        uint64_t p1 = (uint64_t)PHILOX_M1 * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

/* 53-bit open-interval uniform from two 32-bit words */
static double u53(uint32_t lo, uint32_t hi) {
    uint64_t x = ((uint64_t)hi << 32) | lo;
    return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

/* Box-Muller pair from one Philox block */
static void box_muller(const uint32_t w[4], double *z0, double *z1) {
    double u1 = u53(w[0], w[1]);
    double u2 = u53(w[2], w[3]);
    double rr = sqrt(-2.0 * log(u1));
    double th = 6.283185307179586 * u2;
    *z0 = rr * cos(th);
    *z1 = rr * sin(th);
}

/* counter map: ctr = {site, slice k, nucleus, colour pair p}; key = seed */
void orc_rho_normals_site(uint64_t seed, int nucleus, int k, uint32_t s,
                          double z[8]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t p = 0; p < 4; p++) {
        uint32_t ctr[4] = {s, (uint32_t)k, (uint32_t)nucleus, p};
        uint32_t w[4];
        orc_philox4x32(ctr, key, w);
        box_muller(w, &z[2 * p], &z[2 * p + 1]);
    }
}

void orc_rho_normals_slice(uint64_t seed, int nucleus, int k, int S,
                           double *z) {
    for (int s = 0; s < S; s++)
        orc_rho_normals_site(seed, nucleus, k, (uint32_t)s, &z[8 * (size_t)s]);
}

/* restart draws: ctr = {link, restart, 0xFFFF0000 tag, p} */
void orc_newton_restart_normals(uint64_t seed, uint32_t link, uint32_t restart,
                                double alpha[8]) {
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    for (uint32_t p = 0; p < 4; p++) {
        uint32_t ctr[4] = {link, restart, 0xFFFF0000u, p};
        uint32_t w[4];
        orc_philox4x32(ctr, key, w);
        box_muller(w, &alpha[2 * p], &alpha[2 * p + 1]);
    }
}

/* ---- mutation factors ------------------------------------------------------ */
#define ORC_ONES4 1.0, 1.0, 1.0, 1.0
double orc_mut[32] = {ORC_ONES4, ORC_ONES4, ORC_ONES4, ORC_ONES4, ORC_ONES4, ORC_ONES4, ORC_ONES4, ORC_ONES4};

void orc_set_mutation(int k, double factor) {
    if (k >= 0 && k < 32) orc_mut[k] = factor;
}
void orc_clear_mutations(void) {
    for (int k = 0; k < 32; k++) orc_mut[k] = 1.0;
}

/* OpenMP threads of the parallel oracle loops (the CPU baseline times the
 * phases the reference runs serially with one thread) */
#ifdef _OPENMP
#include <omp.h>
#endif
void orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int orc_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
