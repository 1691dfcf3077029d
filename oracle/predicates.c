/*
 * predicates.c - CPU restatement of del3d's filtered exact predicates.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h): the checker the GPU path is
 * compared against, never part of the product.
 *
 * Restates, in C, the semantics of
 *   src/predicates.cpp:22-27   first-stage error-bound constants
 *   src/predicates.cpp:37-39   flip / sign_of
 *   src/predicates.cpp:46-63   orient3d_exact
 *   src/predicates.cpp:76-113  in_sphere_exact
 *   src/predicates.cpp:117-121 lexico_less
 *   src/predicates.cpp:125-147 orient3d (filter, then exact)
 *   src/predicates.cpp:149-204 in_sphere_unperturbed (filter, then exact)
 *   src/predicates.cpp:206-232 in_sphere (symbolic perturbation)
 *   src/expansion.h:33-178     error-free transforms and expansion arithmetic
 * Must be compiled with -ffp-contract=off (the reference's own flag,
 * CMakeLists.txt:81) so that no product is fused into an FMA.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------------- */
/* Scratch arena: expansions are short-lived within one predicate call.    */
/* ---------------------------------------------------------------------- */
typedef struct ArenaBlock {
    struct ArenaBlock* next;
    size_t used, cap;
    double data[];
} ArenaBlock;

static _Thread_local ArenaBlock* g_arena = NULL;

static double* arena_alloc(size_t n) {
    if (n == 0) n = 1;
    ArenaBlock* b = g_arena;
    if (!b || b->used + n > b->cap) {
        size_t cap = n > 65536 ? n : 65536;
        ArenaBlock* nb = (ArenaBlock*)malloc(sizeof(ArenaBlock) + cap * sizeof(double));
        if (!nb) abort();
        nb->next = b;
        nb->used = 0;
        nb->cap = cap;
        g_arena = nb;
        b = nb;
    }
    double* p = b->data + b->used;
    b->used += n;
    return p;
}

static void arena_reset(void) {
    /* keep the newest block, free the rest */
    ArenaBlock* b = g_arena;
    if (!b) return;
    ArenaBlock* rest = b->next;
    b->next = NULL;
    b->used = 0;
    while (rest) { ArenaBlock* n = rest->next; free(rest); rest = n; }
}

/* ---------------------------------------------------------------------- */
/* Error-free transformations (src/expansion.h:33-67).                      */
/* ---------------------------------------------------------------------- */
static inline void eft_sum(double a, double b, double* x, double* err) {
    double s = a + b;
    double bb = s - a;          /* the part of b that made it into s */
    double aa = s - bb;         /* the part of a that made it into s */
    *x = s;
    *err = (a - aa) + (b - bb);
}

static inline void eft_diff(double a, double b, double* x, double* err) {
    eft_sum(a, -b, x, err);     /* negation is exact */
}

static inline void dekker_split(double a, double* hi, double* lo) {
    double t = 134217729.0 * a;     /* 2^27 + 1 */
    double big = t - a;
    *hi = t - big;
    *lo = a - *hi;
}

static inline void eft_product(double a, double b, double* x, double* err) {
    double p = a * b;
    double ah, al, bh, bl;
    dekker_split(a, &ah, &al);
    dekker_split(b, &bh, &bl);
    double r1 = p - (ah * bh);
    double r2 = r1 - (al * bh);
    double r3 = r2 - (ah * bl);
    *x = p;
    *err = (al * bl) - r3;
}

/* ---------------------------------------------------------------------- */
/* Expansions: nonoverlapping components, increasing magnitude, no zeros.   */
/* ---------------------------------------------------------------------- */
typedef struct { int n; double* c; } Exp;

static Exp exp_diff2(double a, double b) {
    double x, e;
    eft_diff(a, b, &x, &e);
    Exp r;
    r.c = arena_alloc(2);
    r.n = 0;
    if (e != 0.0) r.c[r.n++] = e;
    if (x != 0.0) r.c[r.n++] = x;
    return r;
}

/* Add one double into an expansion (the Grow step of src/expansion.h:137-150),
 * writing the result to out (x.n + 1 doubles, not overlapping x). */
static int grow_into(const double* x, int n, double b, double* out) {
    int m = 0;
    double q = b;
    for (int i = 0; i < n; ++i) {
        double s, e;
        eft_sum(q, x[i], &s, &e);
        if (e != 0.0) out[m++] = e;
        q = s;
    }
    if (q != 0.0) out[m++] = q;
    return m;
}

/* x + y: y's components grown into x one at a time, ping-ponging between two
 * buffers (the intermediate expansions are dropped as soon as used). */
static Exp exp_add(Exp x, Exp y) {
    if (y.n == 0) return x;
    const size_t cap = (size_t)x.n + (size_t)y.n;
    double* buf[2] = {arena_alloc(cap), arena_alloc(cap)};
    Exp r = x;
    int w = 0;
    for (int i = 0; i < y.n; ++i) {
        if (y.c[i] == 0.0) continue;
        r.n = grow_into(r.c, r.n, y.c[i], buf[w]);
        r.c = buf[w];
        w ^= 1;
    }
    return r;
}

static Exp exp_neg(Exp x) {
    Exp r;
    r.c = arena_alloc((size_t)x.n);
    r.n = x.n;
    for (int i = 0; i < x.n; ++i) r.c[i] = -x.c[i];
    return r;
}

static Exp exp_sub(Exp x, Exp y) { return exp_add(x, exp_neg(y)); }

/* Scale an expansion by one double (src/expansion.h:155-175) into out
 * (2 * x.n doubles); returns the length. */
static int scale_into(Exp x, double b, double* out) {
    int m = 0;
    if (b == 0.0 || x.n == 0) return 0;
    double q, lo;
    eft_product(x.c[0], b, &q, &lo);
    if (lo != 0.0) out[m++] = lo;
    for (int i = 1; i < x.n; ++i) {
        double ph, pl;
        eft_product(x.c[i], b, &ph, &pl);
        double s1, e1;
        eft_sum(q, pl, &s1, &e1);
        if (e1 != 0.0) out[m++] = e1;
        double s2, e2;
        eft_sum(ph, s1, &s2, &e2);
        if (e2 != 0.0) out[m++] = e2;
        q = s2;
    }
    if (q != 0.0) out[m++] = q;
    return m;
}

/* Product: accumulate x scaled by each component of y (src/expansion.h:127-131).
 * The accumulator ping-pongs between two heap buffers of the largest possible
 * length, so memory stays O(x.n * y.n) however long the operands are; the
 * result is copied into the arena. */
static Exp exp_mul(Exp x, Exp y) {
    Exp r;
    r.n = 0;
    r.c = NULL;
    if (x.n == 0 || y.n == 0) return r;
    const size_t cap = 2 * (size_t)x.n * (size_t)y.n + 1;
    double* buf[3];
    buf[0] = (double*)malloc(cap * sizeof(double));
    buf[1] = (double*)malloc(cap * sizeof(double));
    buf[2] = (double*)malloc(2 * (size_t)x.n * sizeof(double));
    if (!buf[0] || !buf[1] || !buf[2]) abort();
    int n = 0, cur = 0;
    for (int i = 0; i < y.n; ++i) {
        const int m = scale_into(x, y.c[i], buf[2]);
        for (int k = 0; k < m; ++k) {
            n = grow_into(buf[cur], n, buf[2][k], buf[cur ^ 1]);
            cur ^= 1;
        }
    }
    r.n = n;
    r.c = arena_alloc((size_t)n);
    memcpy(r.c, buf[cur], sizeof(double) * (size_t)n);
    free(buf[0]);
    free(buf[1]);
    free(buf[2]);
    return r;
}

static int exp_sign(Exp x) {
    if (x.n == 0) return 0;
    return x.c[x.n - 1] > 0.0 ? 1 : -1;
}

/* ---------------------------------------------------------------------- */
/* Constants and sign helpers (src/predicates.cpp:22-39).                   */
/* ---------------------------------------------------------------------- */
static const double kEps = 1.1102230246251565e-16;                 /* 2^-53 */
#define ORIENT_BOUND ((7.0 + 56.0 * kEps) * kEps)
#define INSPHERE_BOUND ((16.0 + 224.0 * kEps) * kEps)

static inline int sgn(double d) { return d > 0.0 ? 1 : (d < 0.0 ? -1 : 0); }

static _Thread_local int64_t g_stage[3];

/* |a-d, b-d, c-d| exactly; returns its sign (before the interface flip). */
static int orient3d_exact_raw(const double* a, const double* b, const double* c,
                              const double* d) {
    Exp adx = exp_diff2(a[0], d[0]), ady = exp_diff2(a[1], d[1]), adz = exp_diff2(a[2], d[2]);
    Exp bdx = exp_diff2(b[0], d[0]), bdy = exp_diff2(b[1], d[1]), bdz = exp_diff2(b[2], d[2]);
    Exp cdx = exp_diff2(c[0], d[0]), cdy = exp_diff2(c[1], d[1]), cdz = exp_diff2(c[2], d[2]);
    Exp m1 = exp_sub(exp_mul(bdx, cdy), exp_mul(cdx, bdy));
    Exp m2 = exp_sub(exp_mul(cdx, ady), exp_mul(adx, cdy));
    Exp m3 = exp_sub(exp_mul(adx, bdy), exp_mul(bdx, ady));
    Exp det = exp_add(exp_add(exp_mul(adz, m1), exp_mul(bdz, m2)), exp_mul(cdz, m3));
    return exp_sign(det);
}

static int in_sphere_exact_raw(const double* a, const double* b, const double* c,
                               const double* d, const double* e) {
    Exp ax = exp_diff2(a[0], e[0]), ay = exp_diff2(a[1], e[1]), az = exp_diff2(a[2], e[2]);
    Exp bx = exp_diff2(b[0], e[0]), by = exp_diff2(b[1], e[1]), bz = exp_diff2(b[2], e[2]);
    Exp cx = exp_diff2(c[0], e[0]), cy = exp_diff2(c[1], e[1]), cz = exp_diff2(c[2], e[2]);
    Exp dx = exp_diff2(d[0], e[0]), dy = exp_diff2(d[1], e[1]), dz = exp_diff2(d[2], e[2]);

    Exp la = exp_add(exp_add(exp_mul(ax, ax), exp_mul(ay, ay)), exp_mul(az, az));
    Exp lb = exp_add(exp_add(exp_mul(bx, bx), exp_mul(by, by)), exp_mul(bz, bz));
    Exp lc = exp_add(exp_add(exp_mul(cx, cx), exp_mul(cy, cy)), exp_mul(cz, cz));
    Exp ld = exp_add(exp_add(exp_mul(dx, dx), exp_mul(dy, dy)), exp_mul(dz, dz));

    Exp ab = exp_sub(exp_mul(ax, by), exp_mul(bx, ay));
    Exp bc = exp_sub(exp_mul(bx, cy), exp_mul(cx, by));
    Exp cd = exp_sub(exp_mul(cx, dy), exp_mul(dx, cy));
    Exp da = exp_sub(exp_mul(dx, ay), exp_mul(ax, dy));
    Exp ac = exp_sub(exp_mul(ax, cy), exp_mul(cx, ay));
    Exp bd = exp_sub(exp_mul(bx, dy), exp_mul(dx, by));

    Exp abc = exp_add(exp_sub(exp_mul(az, bc), exp_mul(bz, ac)), exp_mul(cz, ab));
    Exp bcd = exp_add(exp_sub(exp_mul(bz, cd), exp_mul(cz, bd)), exp_mul(dz, bc));
    Exp cda = exp_add(exp_add(exp_mul(cz, da), exp_mul(dz, ac)), exp_mul(az, cd));
    Exp dab = exp_add(exp_add(exp_mul(dz, ab), exp_mul(az, bd)), exp_mul(bz, da));

    Exp det = exp_add(exp_sub(exp_mul(ld, abc), exp_mul(lc, dab)),
                      exp_sub(exp_mul(lb, cda), exp_mul(la, bcd)));
    return exp_sign(det);
}

/* Public orient3d: sign of |b-a, c-a, d-a| (the flip of the computed form). */
static int orient3d_impl(const double* a, const double* b, const double* c, const double* d) {
    const double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
    const double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
    const double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];

    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;

    const double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    const double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz)
                      + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz)
                      + (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    const double bound = ORIENT_BOUND * perm;
    if (det > bound || -det > bound) { g_stage[0]++; return -sgn(det); }
    g_stage[1]++;
    return -orient3d_exact_raw(a, b, c, d);
}

static int in_sphere_raw_impl(const double* a, const double* b, const double* c,
                              const double* d, const double* e) {
    const double aex = a[0] - e[0], aey = a[1] - e[1], aez = a[2] - e[2];
    const double bex = b[0] - e[0], bey = b[1] - e[1], bez = b[2] - e[2];
    const double cex = c[0] - e[0], cey = c[1] - e[1], cez = c[2] - e[2];
    const double dex = d[0] - e[0], dey = d[1] - e[1], dez = d[2] - e[2];

    const double ab = aex * bey - bex * aey;
    const double bc = bex * cey - cex * bey;
    const double cd = cex * dey - dex * cey;
    const double da = dex * aey - aex * dey;
    const double ac = aex * cey - cex * aey;
    const double bd = bex * dey - dex * bey;

    const double abc = aez * bc - bez * ac + cez * ab;
    const double bcd = bez * cd - cez * bd + dez * bc;
    const double cda = cez * da + dez * ac + aez * cd;
    const double dab = dez * ab + aez * bd + bez * da;

    const double alift = aex * aex + aey * aey + aez * aez;
    const double blift = bex * bex + bey * bey + bez * bez;
    const double clift = cex * cex + cey * cey + cez * cez;
    const double dlift = dex * dex + dey * dey + dez * dez;

    const double det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);

    const double az_ = fabs(aez), bz_ = fabs(bez), cz_ = fabs(cez), dz_ = fabs(dez);
    const double aexbey = fabs(aex * bey), bexaey = fabs(bex * aey);
    const double bexcey = fabs(bex * cey), cexbey = fabs(cex * bey);
    const double cexdey = fabs(cex * dey), dexcey = fabs(dex * cey);
    const double dexaey = fabs(dex * aey), aexdey = fabs(aex * dey);
    const double aexcey = fabs(aex * cey), cexaey = fabs(cex * aey);
    const double bexdey = fabs(bex * dey), dexbey = fabs(dex * bey);

    const double perm =
        ((cexdey + dexcey) * bz_ + (dexbey + bexdey) * cz_ + (bexcey + cexbey) * dz_) * alift +
        ((dexaey + aexdey) * cz_ + (aexcey + cexaey) * dz_ + (cexdey + dexcey) * az_) * blift +
        ((aexbey + bexaey) * dz_ + (bexdey + dexbey) * az_ + (dexaey + aexdey) * bz_) * clift +
        ((bexcey + cexbey) * az_ + (cexaey + aexcey) * bz_ + (aexbey + bexaey) * cz_) * dlift;

    const double bound = INSPHERE_BOUND * perm;
    if (det > bound || -det > bound) { g_stage[0]++; return -sgn(det); }
    g_stage[1]++;
    return -in_sphere_exact_raw(a, b, c, d, e);
}

static int lex_less(const double* p, const double* q) {
    if (p[0] != q[0]) return p[0] < q[0];
    if (p[1] != q[1]) return p[1] < q[1];
    return p[2] < q[2];
}

/* Stable insertion sort of 5 pointers by lex_less: what libstdc++'s std::sort
 * does for ranges of 16 or fewer elements (a plain insertion sort). */
static void sort5_stable(const double* p[5]) {
    for (int i = 1; i < 5; ++i) {
        const double* v = p[i];
        int j = i;
        while (j > 0 && lex_less(v, p[j - 1])) { p[j] = p[j - 1]; --j; }
        p[j] = v;
    }
}

static int in_sphere_impl(const double* a, const double* b, const double* c,
                          const double* d, const double* e) {
    int s = in_sphere_raw_impl(a, b, c, d, e);
    if (s != 0) return s;
    g_stage[2]++;
    const double* p[5] = {a, b, c, d, e};
    sort5_stable(p);
    for (int i = 4; i > 2; --i) {
        if (p[i] == e) return -1;
        int o;
        if (p[i] == d && (o = orient3d_impl(a, b, c, e)) != 0) return o;
        if (p[i] == c && (o = orient3d_impl(a, b, e, d)) != 0) return o;
        if (p[i] == b && (o = orient3d_impl(a, e, c, d)) != 0) return o;
        if (p[i] == a && (o = orient3d_impl(e, b, c, d)) != 0) return o;
    }
    return -1;
}

int orc_orient3d(const double* a, const double* b, const double* c, const double* d) {
    int r = orient3d_impl(a, b, c, d);
    arena_reset();
    return r;
}

int orc_in_sphere_unperturbed(const double* a, const double* b, const double* c,
                              const double* d, const double* e) {
    int r = in_sphere_raw_impl(a, b, c, d, e);
    arena_reset();
    return r;
}

int orc_in_sphere(const double* a, const double* b, const double* c,
                  const double* d, const double* e) {
    int r = in_sphere_impl(a, b, c, d, e);
    arena_reset();
    return r;
}

void orc_pred_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                    const int32_t* idx5, int64_t n_q, int8_t* sign_out) {
    (void)n_pts;
    g_stage[0] = g_stage[1] = g_stage[2] = 0;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const double* a = xyz + 3 * (int64_t)r[0];
        const double* b = xyz + 3 * (int64_t)r[1];
        const double* c = xyz + 3 * (int64_t)r[2];
        const double* d = xyz + 3 * (int64_t)r[3];
        int s;
        if (op[q] == ORC_OP_ORIENT3D) {
            s = orient3d_impl(a, b, c, d);
        } else {
            const double* e = xyz + 3 * (int64_t)r[4];
            s = op[q] == ORC_OP_INSPHERE ? in_sphere_impl(a, b, c, d, e)
                                         : in_sphere_raw_impl(a, b, c, d, e);
        }
        sign_out[q] = (int8_t)s;
        arena_reset();
    }
}

void orc_last_stage_counts(int64_t out[3]) {
    out[0] = g_stage[0];
    out[1] = g_stage[1];
    out[2] = g_stage[2];
}

/* ====================================================================== */
/* The coplanar predicates (restating src/predicates2.cpp).                */
/*   src/predicates2.cpp:22-26   error-bound constants                     */
/*   src/predicates2.cpp:34-40   orient2d_exact                            */
/*   src/predicates2.cpp:59-101  in_circle_exact (rows pt, rt, qt, v)      */
/*   src/predicates2.cpp:111-180 orient2d .. collinear_position            */
/*   src/predicates2.cpp:182-259 coplanar_in_circle(_unperturbed)          */
/*   src/triangulation.cpp:296-308 the driver's inexact orientation        */
/* ====================================================================== */
#define ORIENT2D_BOUND ((3.0 + 16.0 * kEps) * kEps)
#define INCIRCLE_BOUND ((16.0 + 224.0 * kEps) * kEps)

static int orient2d_impl(double px, double py, double qx, double qy, double rx, double ry) {
    const double qpx = qx - px, qpy = qy - py;
    const double rpx = rx - px, rpy = ry - py;
    const double a = qpx * rpy, b = qpy * rpx;
    const double det = a - b;
    const double bound = ORIENT2D_BOUND * (fabs(a) + fabs(b));
    if (det > bound || -det > bound) { g_stage[0]++; return sgn(det); }
    g_stage[1]++;
    Exp ex = exp_sub(exp_mul(exp_diff2(qx, px), exp_diff2(ry, py)),
                     exp_mul(exp_diff2(qy, py), exp_diff2(rx, px)));
    return exp_sign(ex);
}

static int cop_orient3(const double* p, const double* q, const double* r) {
    int o = orient2d_impl(p[0], p[1], q[0], q[1], r[0], r[1]);
    if (o) return o;
    o = orient2d_impl(p[1], p[2], q[1], q[2], r[1], r[2]);
    if (o) return o;
    return orient2d_impl(p[0], p[2], q[0], q[2], r[0], r[2]);
}

static int cop_orient4(const double* p, const double* q, const double* r, const double* s) {
    int o = orient2d_impl(p[0], p[1], q[0], q[1], r[0], r[1]);
    if (o) return o * orient2d_impl(p[0], p[1], q[0], q[1], s[0], s[1]);
    o = orient2d_impl(p[1], p[2], q[1], q[2], r[1], r[2]);
    if (o) return o * orient2d_impl(p[1], p[2], q[1], q[2], s[1], s[2]);
    o = orient2d_impl(p[0], p[2], q[0], q[2], r[0], r[2]);
    return o * orient2d_impl(p[0], p[2], q[0], q[2], s[0], s[2]);
}

static int cop_collinear(const double* p, const double* q, const double* r) {
    if (orient2d_impl(r[0], r[1], p[0], p[1], q[0], q[1]) != 0) return 0;
    if (orient2d_impl(r[0], r[2], p[0], p[2], q[0], q[2]) != 0) return 0;
    return orient2d_impl(r[1], r[2], p[1], p[2], q[1], q[2]) == 0;
}

static int cop_equal(const double* p, const double* q) {
    return p[0] == q[0] && p[1] == q[1] && p[2] == q[2];
}

/* 0 before p, 1 at p, 2 between, 3 at r, 4 after r */
static int cop_position(const double* p, const double* q, const double* r) {
    if (cop_equal(q, p)) return 1;
    if (cop_equal(q, r)) return 3;
    const int k = p[0] != r[0] ? 0 : (p[1] != r[1] ? 1 : 2);
    const double a = p[k], b = q[k], c = r[k];
    if (a < c) return b < a ? 0 : (b > c ? 4 : 2);
    return b > a ? 0 : (b < c ? 4 : 2);
}

/* The 4x4 determinant of rows (pt, rt, qt, v) lifted, v = pq x pr, its 2x2
 * minors formed as m_ij = a_j0 a_i1 - a_i0 a_j1, in expansion arithmetic. */
static int in_circle_exact_raw(const double* p0, const double* p1, const double* p2,
                               const double* t) {
    Exp a[4][4];
    for (int k = 0; k < 3; ++k) {
        a[0][k] = exp_diff2(p0[k], t[k]);
        a[1][k] = exp_diff2(p2[k], t[k]);
        a[2][k] = exp_diff2(p1[k], t[k]);
    }
    for (int i = 0; i < 3; ++i)
        a[i][3] = exp_add(exp_add(exp_mul(a[i][0], a[i][0]), exp_mul(a[i][1], a[i][1])),
                          exp_mul(a[i][2], a[i][2]));
    Exp pq[3], pr[3];
    for (int k = 0; k < 3; ++k) { pq[k] = exp_diff2(p1[k], p0[k]); pr[k] = exp_diff2(p2[k], p0[k]); }
    a[3][0] = exp_sub(exp_mul(pq[1], pr[2]), exp_mul(pq[2], pr[1]));
    a[3][1] = exp_sub(exp_mul(pq[2], pr[0]), exp_mul(pq[0], pr[2]));
    a[3][2] = exp_sub(exp_mul(pq[0], pr[1]), exp_mul(pq[1], pr[0]));
    a[3][3] = exp_add(exp_add(exp_mul(a[3][0], a[3][0]), exp_mul(a[3][1], a[3][1])),
                      exp_mul(a[3][2], a[3][2]));
#define M2(i, j) exp_sub(exp_mul(a[j][0], a[i][1]), exp_mul(a[i][0], a[j][1]))
    const Exp m01 = M2(0, 1), m02 = M2(0, 2), m03 = M2(0, 3);
    const Exp m12 = M2(1, 2), m13 = M2(1, 3), m23 = M2(2, 3);
#undef M2
#define M3(x, y, z, r0, r1, r2) \
    exp_add(exp_sub(exp_mul(x, a[r0][2]), exp_mul(y, a[r1][2])), exp_mul(z, a[r2][2]))
    const Exp m012 = M3(m12, m02, m01, 0, 1, 2);
    const Exp m013 = M3(m13, m03, m01, 0, 1, 3);
    const Exp m023 = M3(m23, m03, m02, 0, 2, 3);
    const Exp m123 = M3(m23, m13, m12, 1, 2, 3);
#undef M3
    const Exp det = exp_sub(exp_add(exp_sub(exp_mul(m123, a[0][3]), exp_mul(m023, a[1][3])),
                                    exp_mul(m013, a[2][3])),
                            exp_mul(m012, a[3][3]));
    return exp_sign(det);
}

static int in_circle_raw_impl(const double* p0, const double* p1, const double* p2,
                              const double* p) {
    double a[4][4];
    for (int k = 0; k < 3; ++k) {
        a[0][k] = p0[k] - p[k];
        a[2][k] = p1[k] - p[k];
        a[1][k] = p2[k] - p[k];
    }
    a[0][3] = a[0][0] * a[0][0] + a[0][1] * a[0][1] + a[0][2] * a[0][2];
    a[2][3] = a[2][0] * a[2][0] + a[2][1] * a[2][1] + a[2][2] * a[2][2];
    a[1][3] = a[1][0] * a[1][0] + a[1][1] * a[1][1] + a[1][2] * a[1][2];
    double pq[3], pr[3];
    for (int k = 0; k < 3; ++k) { pq[k] = p1[k] - p0[k]; pr[k] = p2[k] - p0[k]; }
    a[3][0] = pq[1] * pr[2] - pq[2] * pr[1];
    a[3][1] = pq[2] * pr[0] - pq[0] * pr[2];
    a[3][2] = pq[0] * pr[1] - pq[1] * pr[0];
    a[3][3] = a[3][0] * a[3][0] + a[3][1] * a[3][1] + a[3][2] * a[3][2];
#define M2(i, j) (a[j][0] * a[i][1] - a[i][0] * a[j][1])
    const double m01 = M2(0, 1), m02 = M2(0, 2), m03 = M2(0, 3);
    const double m12 = M2(1, 2), m13 = M2(1, 3), m23 = M2(2, 3);
#undef M2
    const double m012 = m12 * a[0][2] - m02 * a[1][2] + m01 * a[2][2];
    const double m013 = m13 * a[0][2] - m03 * a[1][2] + m01 * a[3][2];
    const double m023 = m23 * a[0][2] - m03 * a[2][2] + m02 * a[3][2];
    const double m123 = m23 * a[1][2] - m13 * a[2][2] + m12 * a[3][2];
    const double det = m123 * a[0][3] - m023 * a[1][3] + m013 * a[2][3] - m012 * a[3][3];
    double mag = 0.0;
    for (int i = 0; i < 4; ++i) {
        double row = 0.0;
        for (int j = 0; j < 4; ++j) row += fabs(a[i][j]);
        mag = mag == 0.0 ? row : mag * row;
    }
    const double bound = INCIRCLE_BOUND * mag;
    if (det > bound || -det > bound) { g_stage[0]++; return sgn(det); }
    g_stage[1]++;
    return in_circle_exact_raw(p0, p1, p2, p);
}

/* src/triangulation.cpp:296-308: plain double, the driver's association. */
static int inexact_orientation(const double* p, const double* q, const double* r,
                               const double* s) {
    const double a00 = q[0] - p[0], a01 = q[1] - p[1], a02 = q[2] - p[2];
    const double a10 = r[0] - p[0], a11 = r[1] - p[1], a12 = r[2] - p[2];
    const double a20 = s[0] - p[0], a21 = s[1] - p[1], a22 = s[2] - p[2];
    const double m01 = a00 * a11 - a10 * a01;
    const double m02 = a00 * a21 - a20 * a01;
    const double m12 = a10 * a21 - a20 * a11;
    const double det = m01 * a22 - m02 * a12 + m12 * a02;
    return sgn(det);
}

static void sort4_stable(const double* p[4]) {
    for (int i = 1; i < 4; ++i) {
        const double* v = p[i];
        int j = i;
        while (j > 0 && lex_less(v, p[j - 1])) { p[j] = p[j - 1]; --j; }
        p[j] = v;
    }
}

static int in_circle_impl(const double* p0, const double* p1, const double* p2,
                          const double* p) {
    const int s = in_circle_raw_impl(p0, p1, p2, p);
    if (s != 0) return s;
    g_stage[2]++;
    const double* pts[4] = {p0, p1, p2, p};
    sort4_stable(pts);
    const int local = cop_orient3(p0, p1, p2);
    for (int i = 3; i > 0; --i) {
        if (pts[i] == p) return -1;
        int o;
        if (pts[i] == p2 && (o = cop_orient3(p0, p1, p)) != 0) return o * local;
        if (pts[i] == p1 && (o = cop_orient3(p0, p, p2)) != 0) return o * local;
        if (pts[i] == p0 && (o = cop_orient3(p, p1, p2)) != 0) return o * local;
    }
    return -local;
}

int orc_pred2_one(uint8_t op, const double* a, const double* b, const double* c,
                  const double* d) {
    int s;
    switch (op) {
        case ORC_OP2_COPLANAR_ORIENTATION: s = cop_orient3(a, b, c); break;
        case ORC_OP2_COPLANAR_ORIENTATION4: s = cop_orient4(a, b, c, d); break;
        case ORC_OP2_COLLINEAR: s = cop_collinear(a, b, c); break;
        case ORC_OP2_EQUAL: s = cop_equal(a, b); break;
        case ORC_OP2_COLLINEAR_POSITION: s = cop_position(a, b, c); break;
        case ORC_OP2_IN_CIRCLE: s = in_circle_impl(a, b, c, d); break;
        case ORC_OP2_IN_CIRCLE_UNPERTURBED: s = in_circle_raw_impl(a, b, c, d); break;
        case ORC_OP2_ORIENT2D_XY: s = orient2d_impl(a[0], a[1], b[0], b[1], c[0], c[1]); break;
        case ORC_OP2_INEXACT_ORIENTATION: s = inexact_orientation(a, b, c, d); break;
        default: s = 0; break;
    }
    arena_reset();
    return s;
}

int orc_pred2_points(uint8_t op) {
    switch (op) {
        case ORC_OP2_EQUAL: return 2;
        case ORC_OP2_COPLANAR_ORIENTATION4:
        case ORC_OP2_IN_CIRCLE:
        case ORC_OP2_IN_CIRCLE_UNPERTURBED:
        case ORC_OP2_INEXACT_ORIENTATION: return 4;
        default: return 3;
    }
}

void orc_pred2_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                     const int32_t* idx5, int64_t n_q, int8_t* sign_out) {
    (void)n_pts;
    g_stage[0] = g_stage[1] = g_stage[2] = 0;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const int np = orc_pred2_points(op[q]);
        const double* p[4];
        for (int k = 0; k < 4; ++k) p[k] = k < np ? xyz + 3 * (int64_t)r[k] : xyz;
        sign_out[q] = (int8_t)orc_pred2_one(op[q], p[0], p[1], p[2], p[3]);
    }
}

/* ====================================================================== */
/* Circumcentres (restating src/triangulation.cpp:805-930 with             */
/* src/interval.h:27-186).                                                 */
/* ====================================================================== */
#include <fenv.h>

/* The exact closed form: numerator and denominator as expansions, one
 * rounding per coordinate at the end. */
static void circumcenter_exact(const double* p0, const double* p1, const double* p2,
                               const double* p3, double out[3]) {
    Exp a[3], b[3], c[3];
    for (int k = 0; k < 3; ++k) {
        a[k] = exp_diff2(p1[k], p0[k]);
        b[k] = exp_diff2(p2[k], p0[k]);
        c[k] = exp_diff2(p3[k], p0[k]);
    }
    /* u x v, coordinate k = u[k+1] v[k+2] - u[k+2] v[k+1] */
    Exp bc[3], ca[3], ab[3];
    for (int k = 0; k < 3; ++k) {
        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
        bc[k] = exp_sub(exp_mul(b[k1], c[k2]), exp_mul(b[k2], c[k1]));
        ca[k] = exp_sub(exp_mul(c[k1], a[k2]), exp_mul(c[k2], a[k1]));
        ab[k] = exp_sub(exp_mul(a[k1], b[k2]), exp_mul(a[k2], b[k1]));
    }
    const Exp a2 = exp_add(exp_add(exp_mul(a[0], a[0]), exp_mul(a[1], a[1])), exp_mul(a[2], a[2]));
    const Exp b2 = exp_add(exp_add(exp_mul(b[0], b[0]), exp_mul(b[1], b[1])), exp_mul(b[2], b[2]));
    const Exp c2 = exp_add(exp_add(exp_mul(c[0], c[0]), exp_mul(c[1], c[1])), exp_mul(c[2], c[2]));
    double num[3];
    for (int k = 0; k < 3; ++k) {
        const Exp n = exp_add(exp_add(exp_mul(a2, bc[k]), exp_mul(b2, ca[k])), exp_mul(c2, ab[k]));
        double s = 0.0;
        for (int i = 0; i < n.n; ++i) s += n.c[i];
        num[k] = s;
    }
    Exp two;
    two.n = 1;
    two.c = arena_alloc(1);
    two.c[0] = 2.0;
    const Exp den = exp_mul(exp_add(exp_add(exp_mul(a[0], bc[0]), exp_mul(a[1], bc[1])),
                                    exp_mul(a[2], bc[2])),
                            two);
    if (den.n == 0) {
        out[0] = p0[0]; out[1] = p0[1]; out[2] = p0[2];
        return;
    }
    double d = 0.0;
    for (int i = 0; i < den.n; ++i) d += den.c[i];
    for (int k = 0; k < 3; ++k) out[k] = p0[k] + num[k] / d;
}

/* Intervals under FE_UPWARD, the lower bound stored negated; every
 * operation forced through memory so nothing is folded or reassociated. */
typedef struct { double ni, s; } Iv;
static double opq(double x) { volatile double v = x; return v; }
static double iv_lo(Iv a) { return -a.ni; }
static Iv iv_make(double i, double s) { Iv r; r.ni = -i; r.s = s; return r; }
static Iv iv_pt(double d) { return iv_make(d, d); }
static double up_mul(double a, double b) { return opq(opq(a) * opq(b)); }
static double up_div(double a, double b) { return opq(opq(a) / opq(b)); }
static double dmax(double a, double b) { return a < b ? b : a; }

static Iv iv_add(Iv a, Iv b) {
    return iv_make(-opq(opq(-iv_lo(a)) + opq(-iv_lo(b))), opq(opq(a.s) + opq(b.s)));
}
static Iv iv_sub(Iv a, Iv b) {
    return iv_make(-opq(opq(-iv_lo(a)) + opq(b.s)), opq(opq(a.s) + opq(-iv_lo(b))));
}
static Iv iv_mul(Iv a, Iv b) {
    if (iv_lo(a) >= 0.0) {
        double aa = iv_lo(a), bb = a.s;
        if (bb <= 0.0) return iv_pt(0.0);
        if (iv_lo(b) < 0.0) {
            aa = bb;
            if (b.s < 0.0) bb = iv_lo(a);
        }
        const double r = (b.s == 0) ? 0.0 : up_mul(bb, b.s);
        return iv_make(-up_mul(aa, -iv_lo(b)), r);
    }
    if (a.s <= 0.0) {
        double aa = a.s, bb = iv_lo(a);
        if (iv_lo(b) < 0.0) {
            aa = bb;
            if (b.s <= 0.0) bb = a.s;
        } else if (b.s <= 0) {
            return iv_pt(0.0);
        }
        return iv_make(-up_mul(-bb, b.s), up_mul(-aa, -iv_lo(b)));
    }
    if (iv_lo(b) >= 0.0) {
        if (b.s <= 0.0) return iv_pt(0.0);
        return iv_make(-up_mul(-iv_lo(a), b.s), up_mul(a.s, b.s));
    }
    if (b.s <= 0.0) return iv_make(-up_mul(a.s, -iv_lo(b)), up_mul(-iv_lo(a), -iv_lo(b)));
    const double t1 = up_mul(-iv_lo(a), b.s), t2 = up_mul(a.s, -iv_lo(b));
    const double t3 = up_mul(-iv_lo(a), -iv_lo(b)), t4 = up_mul(a.s, b.s);
    return iv_make(-dmax(t1, t2), dmax(t3, t4));
}
static Iv iv_div(Iv a, Iv b) {
    if (iv_lo(b) > 0.0) {
        double aa = b.s, bb = iv_lo(b);
        if (iv_lo(a) < 0.0) {
            aa = bb;
            if (a.s < 0.0) bb = b.s;
        }
        return iv_make(-up_div(-iv_lo(a), aa), up_div(a.s, bb));
    }
    if (b.s < 0.0) {
        double aa = b.s, bb = iv_lo(b);
        if (iv_lo(a) < 0.0) {
            bb = aa;
            if (a.s < 0.0) aa = iv_lo(b);
        }
        return iv_make(-up_div(a.s, -aa), up_div(iv_lo(a), bb));
    }
    return iv_make(-INFINITY, INFINITY);
}
static Iv iv_sq(Iv d) {
    if (iv_lo(d) >= 0.0) return iv_make(-up_mul(-iv_lo(d), iv_lo(d)), up_mul(d.s, d.s));
    if (d.s <= 0.0) return iv_make(-up_mul(d.s, -d.s), up_mul(-iv_lo(d), -iv_lo(d)));
    const double m = dmax(-iv_lo(d), d.s);
    return iv_make(0.0, up_mul(m, m));
}
static Iv iv_det3(Iv a00, Iv a01, Iv a02, Iv a10, Iv a11, Iv a12, Iv a20, Iv a21, Iv a22) {
    const Iv m01 = iv_sub(iv_mul(a00, a11), iv_mul(a10, a01));
    const Iv m02 = iv_sub(iv_mul(a00, a21), iv_mul(a20, a01));
    const Iv m12 = iv_sub(iv_mul(a10, a21), iv_mul(a20, a11));
    return iv_add(iv_sub(iv_mul(m01, a22), iv_mul(m02, a12)), iv_mul(m12, a02));
}

static void circumcenter_lazy(const double* p, const double* q, const double* r, const double* s,
                              double out[3]) {
    int need[3] = {0, 0, 0};
    const int saved = fegetround();
    fesetround(FE_UPWARD);
    {
        const Iv px = iv_pt(p[0]), py = iv_pt(p[1]), pz = iv_pt(p[2]);
        const Iv qpx = iv_sub(iv_pt(q[0]), px), qpy = iv_sub(iv_pt(q[1]), py), qpz = iv_sub(iv_pt(q[2]), pz);
        const Iv qp2 = iv_add(iv_add(iv_sq(qpx), iv_sq(qpy)), iv_sq(qpz));
        const Iv rpx = iv_sub(iv_pt(r[0]), px), rpy = iv_sub(iv_pt(r[1]), py), rpz = iv_sub(iv_pt(r[2]), pz);
        const Iv rp2 = iv_add(iv_add(iv_sq(rpx), iv_sq(rpy)), iv_sq(rpz));
        const Iv spx = iv_sub(iv_pt(s[0]), px), spy = iv_sub(iv_pt(s[1]), py), spz = iv_sub(iv_pt(s[2]), pz);
        const Iv sp2 = iv_add(iv_add(iv_sq(spx), iv_sq(spy)), iv_sq(spz));
        const Iv nx = iv_det3(qpy, qpz, qp2, rpy, rpz, rp2, spy, spz, sp2);
        const Iv ny = iv_det3(qpx, qpz, qp2, rpx, rpz, rp2, spx, spz, sp2);
        const Iv nz = iv_det3(qpx, qpy, qp2, rpx, rpy, rp2, spx, spy, sp2);
        const Iv den = iv_det3(qpx, qpy, qpz, rpx, rpy, rpz, spx, spy, spz);
        const Iv inv = iv_div(iv_pt(1.0), iv_mul(iv_pt(2.0), den));
        const Iv c[3] = {iv_add(px, iv_mul(nx, inv)), iv_sub(py, iv_mul(ny, inv)),
                         iv_add(pz, iv_mul(nz, inv))};
        const double prec = 0.00001;
        for (int i = 0; i < 3; ++i) {
            const double lo = iv_lo(c[i]), hi = c[i].s;
            const double mag = dmax(fabs(lo), fabs(hi));
            const double radius = opq(opq(opq(hi) - opq(lo)) / 2);
            if (hi == lo) out[i] = lo;
            else if (mag == 0 || radius < opq(prec * mag)) out[i] = opq(opq(hi + lo) * 0.5);
            else need[i] = 1;
        }
    }
    fesetround(saved);
    if (need[0] || need[1] || need[2]) {
        double ex[3];
        circumcenter_exact(p, q, r, s, ex);
        for (int i = 0; i < 3; ++i)
            if (need[i]) out[i] = ex[i];
    }
}

void orc_circumcenters(const double* xyz, const int32_t* cells, int64_t n, int lazy, double* out) {
    for (int64_t c = 0; c < n; ++c) {
        const int32_t* v = cells + 4 * c;
        const double* p[4];
        for (int k = 0; k < 4; ++k) p[k] = xyz + 3 * (int64_t)v[k];
        if (lazy) circumcenter_lazy(p[0], p[1], p[2], p[3], out + 3 * c);
        else circumcenter_exact(p[0], p[1], p[2], p[3], out + 3 * c);
        arena_reset();
    }
}
