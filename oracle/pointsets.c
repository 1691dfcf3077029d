/*
 * pointsets.c - deterministic point sets of the reference's tests.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h). Regenerates the inputs behind the
 * reference's pinned hashes, so the same mapping onto doubles is kept:
 *   tests/pointsets.h:36-45   Rng: std::mt19937 words; u01 = w * 2^-32;
 *                             range(lo,hi) = lo + (hi-lo)*u01; below(n) = w % n
 *   tests/pointsets.h:48      snap(v,q) = floor(v*q + 0.5)/q
 *   tests/pointsets.h:51-126  cylinder, sphere, torus, box_shell, lattice,
 *                             random_cube
 * mt19937 itself is the published Matsumoto-Nishimura generator (state 624,
 * the standard's 1812433253 seeding recurrence and tempering), written here
 * from that published specification. Built with -ffp-contract=off; the
 * transcendental calls go to the C library's cos/sin/sqrt/floor, the same
 * ones the reference's tests call.
 */
#include "oracle.h"

#include <math.h>

typedef struct { uint32_t s[624]; int i; } MT;

static void mt_seed(MT* m, uint32_t seed) {
    m->s[0] = seed;
    for (int k = 1; k < 624; ++k)
        m->s[k] = 1812433253u * (m->s[k - 1] ^ (m->s[k - 1] >> 30)) + (uint32_t)k;
    m->i = 624;
}

static uint32_t mt_next(MT* m) {
    if (m->i >= 624) {
        for (int k = 0; k < 624; ++k) {
            uint32_t y = (m->s[k] & 0x80000000u) | (m->s[(k + 1) % 624] & 0x7fffffffu);
            uint32_t v = m->s[(k + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            m->s[k] = v;
        }
        m->i = 0;
    }
    uint32_t y = m->s[m->i++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

static double u01(MT* m) { return (double)mt_next(m) * (1.0 / 4294967296.0); }
static double rng_range(MT* m, double lo, double hi) { return lo + (hi - lo) * u01(m); }
static double snapq(double v, int q) { return floor(v * q + 0.5) / q; }

static const double kTwoPi = 6.283185307179586;

void orc_mt19937_words(uint32_t seed, int64_t count, uint32_t* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = mt_next(&m);
}

void orc_rng_below(uint32_t seed, uint64_t n, int64_t count, int64_t* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = (int64_t)((uint64_t)mt_next(&m) % n);
}

/* n points; writes 3n doubles */
int64_t orc_pts_cylinder(int64_t n, int q, uint32_t seed, double* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < n; ++i) {
        const double a = kTwoPi * u01(&m);
        out[3 * i + 0] = snapq(2.0 * cos(a), q);
        out[3 * i + 1] = snapq(2.0 * sin(a), q);
        out[3 * i + 2] = snapq(rng_range(&m, -1.5, 1.5), q);
    }
    return n;
}

/* n points; writes 3n doubles */
int64_t orc_pts_sphere(int64_t n, int q, uint32_t seed, double* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < n; ++i) {
        const double z = rng_range(&m, -1.0, 1.0);
        const double r = sqrt(1.0 - z * z);
        const double t = kTwoPi * u01(&m);
        out[3 * i + 0] = snapq(2.0 * r * cos(t), q);
        out[3 * i + 1] = snapq(2.0 * r * sin(t), q);
        out[3 * i + 2] = snapq(2.0 * z, q);
    }
    return n;
}

/* n points; writes 3n doubles */
int64_t orc_pts_torus(int64_t n, int q, uint32_t seed, double* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < n; ++i) {
        const double u = kTwoPi * u01(&m);
        const double v = kTwoPi * u01(&m);
        const double R = 2.0 + 0.75 * cos(v);
        out[3 * i + 0] = snapq(R * cos(u), q);
        out[3 * i + 1] = snapq(R * sin(u), q);
        out[3 * i + 2] = snapq(0.75 * sin(v), q);
    }
    return n;
}

/* (m+1)^3 - (m-1)^3 points for m >= 1 (the lattice surface); out may be NULL
 * to only count. */
int64_t orc_pts_box_shell(int m, double* out) {
    int64_t k = 0;
    for (int i = 0; i <= m; ++i)
        for (int j = 0; j <= m; ++j)
            for (int l = 0; l <= m; ++l)
                if (i == 0 || i == m || j == 0 || j == m || l == 0 || l == m) {
                    if (out) {
                        out[3 * k + 0] = (double)i;
                        out[3 * k + 1] = (double)j;
                        out[3 * k + 2] = (double)l;
                    }
                    ++k;
                }
    return k;
}

/* nx*ny*nz points, index (i*ny + j)*nz + k */
int64_t orc_pts_lattice(int nx, int ny, int nz, double* out) {
    int64_t k = 0;
    for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
            for (int l = 0; l < nz; ++l) {
                out[3 * k + 0] = (double)i;
                out[3 * k + 1] = (double)j;
                out[3 * k + 2] = (double)l;
                ++k;
            }
    return k;
}

/* n points uniform in [-1,1)^3 */
int64_t orc_pts_random_cube(int64_t n, uint32_t seed, double* out) {
    MT m;
    mt_seed(&m, seed);
    for (int64_t i = 0; i < n; ++i) {
        out[3 * i + 0] = rng_range(&m, -1.0, 1.0);
        out[3 * i + 1] = rng_range(&m, -1.0, 1.0);
        out[3 * i + 2] = rng_range(&m, -1.0, 1.0);
    }
    return n;
}
