/*
 * spatial_sort.c - CPU restatement of del3d's spatial sort.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates the semantics of
 *   src/rand48.h:26-50     48-bit LCG (x <- 0x5DEECE66D x + 0xB mod 2^48, out x>>17)
 *   src/rand48.h:113-127   unbiased bucket reduction (the only branch reachable
 *                          with a 31-bit engine and int indices)
 *   src/rand48.h:132-153   ShuffleRng + forward Fisher-Yates
 *   src/spatial_sort.cpp:17-36   thresholds and the axis comparator
 *   src/spatial_sort.cpp:40-51   Lomuto partition
 *   src/spatial_sort.cpp:60-79   middle-pivot quickselect, median split
 *   src/spatial_sort.cpp:95-121  Hilbert recursion table
 *   src/spatial_sort.cpp:130-151 multiscale pass and entry point
 */
#include "oracle.h"

#include <stdlib.h>

/* ---- rand48 ----------------------------------------------------------- */
typedef struct { uint64_t x; } R48;

static void r48_seed(R48* g, uint32_t v) {
    g->x = (((uint64_t)v << 16) | 0x330Eull) & ((1ull << 48) - 1);
}

static uint32_t r48_next(R48* g) {
    g->x = (0x5DEECE66Dull * g->x + 0xBull) & ((1ull << 48) - 1);
    return (uint32_t)(g->x >> 17);
}

/* Uniform integer in [0, k] for 1 <= k < 2^31 - 1: bucket the 31-bit draw
 * into k+1 buckets of size floor(2^31/(k+1)) and redraw on the partial one. */
static int64_t r48_uniform(R48* g, uint32_t k, int64_t* rejects) {
    const uint32_t bucket = 0x80000000u / (k + 1u);
    for (;;) {
        uint32_t r = r48_next(g) / bucket;
        if (r <= k) return r;
        if (rejects) (*rejects)++;
    }
}

void orc_rand48_draws(uint32_t seed, int64_t count, uint32_t* out) {
    R48 g;
    r48_seed(&g, seed);
    for (int64_t i = 0; i < count; ++i) out[i] = r48_next(&g);
}

static void shuffle_impl(int64_t n, int32_t* a, int64_t* rejects) {
    R48 g;
    r48_seed(&g, 1);
    if (n <= 0) return;
    for (int64_t k = 1; k < n; ++k) {
        int64_t j = r48_uniform(&g, (uint32_t)k, rejects);
        int32_t t = a[k];
        a[k] = a[j];
        a[j] = t;
    }
}

void orc_shuffle(int64_t n, int32_t* order) { shuffle_impl(n, order, NULL); }

int64_t orc_shuffle_rejections(int64_t n) {
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!tmp) abort();
    for (int64_t i = 0; i < n; ++i) tmp[i] = (int32_t)i;
    int64_t rej = 0;
    shuffle_impl(n, tmp, &rej);
    free(tmp);
    return rej;
}

/* ---- Hilbert median sort ---------------------------------------------- */
typedef struct {
    const double* xyz;
    int32_t* a;
    int64_t visits;
} Ctx;

static inline int cmp_less(const Ctx* c, int axis, int up, int32_t i, int32_t j) {
    const double ci = c->xyz[3 * (int64_t)i + axis];
    const double cj = c->xyz[3 * (int64_t)j + axis];
    return up ? (cj < ci) : (ci < cj);
}

static inline void swp(int32_t* a, int64_t i, int64_t j) {
    int32_t t = a[i];
    a[i] = a[j];
    a[j] = t;
}

/* Lomuto around the element at pv; r is the LAST index of the range. */
static int64_t lomuto(Ctx* c, int64_t l, int64_t r, int64_t pv, int axis, int up) {
    int32_t* a = c->a;
    swp(a, pv, r);
    int64_t res = l;
    for (int64_t i = l; i < r; ++i) {
        if (cmp_less(c, axis, up, a[i], a[r])) {
            swp(a, res, i);
            ++res;
        }
    }
    c->visits += r - l;
    swp(a, r, res);
    return res;
}

static void select_nth(Ctx* c, int64_t l, int64_t nth, int64_t end, int axis, int up) {
    if (l == end) return;
    int64_t r = end - 1;
    if (l == r) return;
    for (;;) {
        int64_t pv = l + (r - l) / 2;
        int64_t p = lomuto(c, l, r, pv, axis, up);
        if (p == nth) return;
        if (nth < p) r = p - 1;
        else l = p + 1;
    }
}

static int64_t median_split(Ctx* c, int64_t b, int64_t e, int axis, int up) {
    if (b >= e) return b;
    int64_t m = b + (e - b) / 2;
    select_nth(c, b, m, e, axis, up);
    return m;
}

static void hilbert(Ctx* c, int64_t b, int64_t e, int x, int ux, int uy, int uz) {
    const int y = (x + 1) % 3, z = (x + 2) % 3;
    if (e - b <= 8) return;
    const int64_t m0 = b, m8 = e;
    const int64_t m4 = median_split(c, m0, m8, x, ux);
    const int64_t m2 = median_split(c, m0, m4, y, uy);
    const int64_t m1 = median_split(c, m0, m2, z, uz);
    const int64_t m3 = median_split(c, m2, m4, z, !uz);
    const int64_t m6 = median_split(c, m4, m8, y, !uy);
    const int64_t m5 = median_split(c, m4, m6, z, uz);
    const int64_t m7 = median_split(c, m6, m8, z, !uz);
    hilbert(c, m0, m1, z, uz, ux, uy);
    hilbert(c, m1, m2, y, uy, uz, ux);
    hilbert(c, m2, m3, y, uy, uz, ux);
    hilbert(c, m3, m4, x, ux, !uy, !uz);
    hilbert(c, m4, m5, x, ux, !uy, !uz);
    hilbert(c, m5, m6, y, !uy, uz, !ux);
    hilbert(c, m6, m7, y, !uy, uz, !ux);
    hilbert(c, m7, m8, z, !uz, !ux, uy);
}

static void multiscale(Ctx* c, int64_t b, int64_t e) {
    int64_t mid = b;
    if (e - b >= 64) {
        mid = b + (int64_t)((double)(e - b) * 0.125);
        multiscale(c, b, mid);
    }
    hilbert(c, mid, e, 0, 0, 0, 0);
}

static int64_t sort_impl(const double* xyz, int64_t n, int32_t* order) {
    if (n < 2) return 0;
    shuffle_impl(n, order, NULL);
    Ctx c = {xyz, order, 0};
    multiscale(&c, 0, n);
    return c.visits;
}

void orc_spatial_sort(const double* xyz, int64_t n, int32_t* order) {
    (void)sort_impl(xyz, n, order);
}

int64_t orc_spatial_sort_visits(const double* xyz, int64_t n) {
    int32_t* tmp = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    if (!tmp) abort();
    for (int64_t i = 0; i < n; ++i) tmp[i] = (int32_t)i;
    int64_t v = sort_impl(xyz, n, tmp);
    free(tmp);
    return v;
}

uint64_t orc_hash_order(const int32_t* order, int64_t n) {
    uint64_t h = 1469598103934665603ull;
    for (int64_t i = 0; i < n; ++i) {
        /* the index widened to 64 bits, then its 4 low bytes */
        uint64_t v = (uint64_t)(int64_t)order[i];
        for (int b = 0; b < 4; ++b) {
            h ^= (v >> (8 * b)) & 0xffull;
            h *= 1099511628211ull;
        }
    }
    return h;
}
