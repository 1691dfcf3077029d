/*
 * exact_int.c - an independent big-integer oracle for the 3D predicates.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows the method of the reference's own test oracle, tests/exact.h:161-291:
 * every finite double is m * 2^e with m an integer, so the coordinates of one
 * predicate call are rewritten as integers on their common smallest scale
 * (which multiplies the determinant by a positive power of two and so keeps
 * its sign), and the determinant is then evaluated in exact integer
 * arithmetic with no floating point anywhere in the decision. It shares no
 * arithmetic with predicates.c, which is what makes it a check on it.
 *
 * Sign conventions (tests/exact.h:242-291, pinned by the known answers of
 * tests/test_predicates.cpp:64-76):
 *   orient3d = sign |b-a, c-a, d-a|
 *   in_sphere_raw = -sign det4_lifted(a-e, b-e, c-e, d-e)
 *   in_sphere = the perturbation rule of include/del3d/predicates.h:28-48.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* 32-bit limbs, sign-magnitude, fixed capacity. The widest product is the
 * coplanar in-circle's degree-7 term: 7 factors of up to 2099+53 bits each
 * need < 15100 bits; 512 limbs hold 16384. */
#define BI_CAP 512

typedef struct {
    int sg;                 /* -1, 0, +1 */
    int n;                  /* limbs in use (no leading zero limbs) */
    uint32_t d[BI_CAP];
} Big;

static void big_copy(Big* r, const Big* a) {
    if (r == a) return;
    r->sg = a->sg;
    r->n = a->n;
    memcpy(r->d, a->d, sizeof(uint32_t) * (size_t)a->n);
}

static void big_trim(Big* x) {
    while (x->n > 0 && x->d[x->n - 1] == 0) x->n--;
    if (x->n == 0) x->sg = 0;
}

static void big_from_u64_shift(Big* x, int sg, uint64_t m, unsigned shift) {
    x->sg = 0;
    x->n = 0;
    if (m == 0) return;
    unsigned w = shift / 32, r = shift % 32;
    if (w + 4 > BI_CAP) abort();
    for (unsigned i = 0; i < w; ++i) x->d[i] = 0;
    /* m << r spans at most 96 bits */
    uint64_t lo = m << r;
    uint64_t hi = r ? (m >> (64 - r)) : 0;
    x->d[w] = (uint32_t)lo;
    x->d[w + 1] = (uint32_t)(lo >> 32);
    x->d[w + 2] = (uint32_t)hi;
    x->n = (int)w + 3;
    x->sg = sg;
    big_trim(x);
}

static int mag_cmp(const Big* a, const Big* b) {
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    for (int i = a->n - 1; i >= 0; --i)
        if (a->d[i] != b->d[i]) return a->d[i] < b->d[i] ? -1 : 1;
    return 0;
}

/* r = |a| + |b| */
static void mag_add(Big* r, const Big* a, const Big* b) {
    int n = a->n > b->n ? a->n : b->n;
    if (n + 1 > BI_CAP) abort();
    uint64_t carry = 0;
    for (int i = 0; i < n; ++i) {
        uint64_t s = carry + (i < a->n ? a->d[i] : 0u) + (i < b->n ? b->d[i] : 0u);
        r->d[i] = (uint32_t)s;
        carry = s >> 32;
    }
    r->d[n] = (uint32_t)carry;
    r->n = n + 1;
}

/* r = |a| - |b|, requires |a| >= |b| */
static void mag_sub(Big* r, const Big* a, const Big* b) {
    int64_t borrow = 0;
    for (int i = 0; i < a->n; ++i) {
        int64_t s = (int64_t)a->d[i] - borrow - (i < b->n ? (int64_t)b->d[i] : 0);
        if (s < 0) { s += ((int64_t)1 << 32); borrow = 1; } else borrow = 0;
        r->d[i] = (uint32_t)s;
    }
    r->n = a->n;
}

static void big_add(Big* r, const Big* a, const Big* b) {
    Big t;
    if (a->sg == 0) { big_copy(r, b); return; }
    if (b->sg == 0) { big_copy(r, a); return; }
    if (a->sg == b->sg) {
        mag_add(&t, a, b);
        t.sg = a->sg;
    } else {
        int c = mag_cmp(a, b);
        if (c == 0) { r->sg = 0; r->n = 0; return; }
        if (c > 0) { mag_sub(&t, a, b); t.sg = a->sg; }
        else       { mag_sub(&t, b, a); t.sg = b->sg; }
    }
    big_trim(&t);
    big_copy(r, &t);
}

static void big_neg(Big* r, const Big* a) { big_copy(r, a); r->sg = -r->sg; }

static void big_sub(Big* r, const Big* a, const Big* b) {
    Big nb;
    big_neg(&nb, b);
    big_add(r, a, &nb);
}

static void big_mul(Big* r, const Big* a, const Big* b) {
    Big t;
    if (a->sg == 0 || b->sg == 0) { r->sg = 0; r->n = 0; return; }
    int n = a->n + b->n;
    if (n > BI_CAP) abort();
    memset(t.d, 0, sizeof(uint32_t) * (size_t)n);
    for (int i = 0; i < a->n; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < b->n; ++j) {
            uint64_t cur = (uint64_t)t.d[i + j] + (uint64_t)a->d[i] * b->d[j] + carry;
            t.d[i + j] = (uint32_t)cur;
            carry = cur >> 32;
        }
        int k = i + b->n;
        while (carry) {
            uint64_t cur = (uint64_t)t.d[k] + carry;
            t.d[k] = (uint32_t)cur;
            carry = cur >> 32;
            ++k;
        }
    }
    t.n = n;
    t.sg = a->sg * b->sg;
    big_trim(&t);
    big_copy(r, &t);
}

/* ---- exact conversion onto a common power-of-two scale ----------------- */
/* x = sg * m * 2^e with m odd (or x == 0). */
static void split_double(double x, int* sg, uint64_t* m, int* e) {
    if (x == 0.0) { *sg = 0; *m = 0; *e = 0; return; }
    int be;
    double f = frexp(fabs(x), &be);               /* |x| = f * 2^be, f in [0.5,1) */
    uint64_t mm = (uint64_t)ldexp(f, 53);         /* exact */
    int ee = be - 53;
    while ((mm & 1u) == 0) { mm >>= 1; ++ee; }
    *sg = x < 0 ? -1 : 1;
    *m = mm;
    *e = ee;
}

typedef struct { int have; int emin; } Frame;

static void frame_add(Frame* f, double x) {
    int sg, e; uint64_t m;
    split_double(x, &sg, &m, &e);
    if (sg == 0) return;
    if (!f->have || e < f->emin) { f->emin = e; f->have = 1; }
}

static void frame_conv(const Frame* f, double x, Big* out) {
    int sg, e; uint64_t m;
    split_double(x, &sg, &m, &e);
    if (sg == 0) { out->sg = 0; out->n = 0; return; }
    big_from_u64_shift(out, sg, m, (unsigned)(e - f->emin));
}

typedef struct { Big x, y, z; } V3;

static void v3_sub(V3* r, const V3* a, const V3* b) {
    big_sub(&r->x, &a->x, &b->x);
    big_sub(&r->y, &a->y, &b->y);
    big_sub(&r->z, &a->z, &b->z);
}

/* the 2x2 minor p*s - q*r */
static void minor2(Big* out, const Big* p, const Big* s, const Big* q, const Big* r) {
    Big t1, t2;
    big_mul(&t1, p, s);
    big_mul(&t2, q, r);
    big_sub(out, &t1, &t2);
}

/* det [a;b;c] by cofactors along the first row */
static void det3(Big* out, const V3* a, const V3* b, const V3* c) {
    Big m, t, acc;
    minor2(&m, &b->y, &c->z, &b->z, &c->y);
    big_mul(&acc, &a->x, &m);
    minor2(&m, &b->x, &c->z, &b->z, &c->x);
    big_mul(&t, &a->y, &m);
    big_sub(&acc, &acc, &t);
    minor2(&m, &b->x, &c->y, &b->y, &c->x);
    big_mul(&t, &a->z, &m);
    big_add(out, &acc, &t);
}

static void norm2(Big* out, const V3* a) {
    Big t, acc;
    big_mul(&acc, &a->x, &a->x);
    big_mul(&t, &a->y, &a->y);
    big_add(&acc, &acc, &t);
    big_mul(&t, &a->z, &a->z);
    big_add(out, &acc, &t);
}

/* 4x4 determinant with rows (p, |p|^2), expanded along the lifted column. */
static void det4_lifted(Big* out, const V3* a, const V3* b, const V3* c, const V3* d) {
    Big n, m, t, acc;
    norm2(&n, d); det3(&m, a, b, c); big_mul(&acc, &n, &m);
    norm2(&n, c); det3(&m, a, b, d); big_mul(&t, &n, &m); big_sub(&acc, &acc, &t);
    norm2(&n, b); det3(&m, a, c, d); big_mul(&t, &n, &m); big_add(&acc, &acc, &t);
    norm2(&n, a); det3(&m, b, c, d); big_mul(&t, &n, &m); big_sub(out, &acc, &t);
}

static void to_v3(const Frame* f, const double* p, V3* v) {
    frame_conv(f, p[0], &v->x);
    frame_conv(f, p[1], &v->y);
    frame_conv(f, p[2], &v->z);
}

int orc_int_orient3d(const double* a, const double* b, const double* c, const double* d) {
    Frame f = {0, 0};
    for (int k = 0; k < 3; ++k) { frame_add(&f, a[k]); frame_add(&f, b[k]); frame_add(&f, c[k]); frame_add(&f, d[k]); }
    V3* A = (V3*)malloc(sizeof(V3) * 7);
    if (!A) abort();
    V3 *B = A + 1, *C = A + 2, *D = A + 3, *u = A + 4, *v = A + 5, *w = A + 6;
    to_v3(&f, a, A); to_v3(&f, b, B); to_v3(&f, c, C); to_v3(&f, d, D);
    v3_sub(u, B, A); v3_sub(v, C, A); v3_sub(w, D, A);
    Big* det = (Big*)malloc(sizeof(Big));
    if (!det) abort();
    det3(det, u, v, w);
    int s = det->sg;
    free(det);
    free(A);
    return s;
}

int orc_int_in_sphere_raw(const double* a, const double* b, const double* c,
                          const double* d, const double* e) {
    Frame f = {0, 0};
    for (int k = 0; k < 3; ++k) {
        frame_add(&f, a[k]); frame_add(&f, b[k]); frame_add(&f, c[k]);
        frame_add(&f, d[k]); frame_add(&f, e[k]);
    }
    V3* V = (V3*)malloc(sizeof(V3) * 9);
    if (!V) abort();
    V3 *E = V, *P = V + 1, *A = V + 2, *B = V + 3, *C = V + 4, *D = V + 5;
    to_v3(&f, e, E);
    to_v3(&f, a, P); v3_sub(A, P, E);
    to_v3(&f, b, P); v3_sub(B, P, E);
    to_v3(&f, c, P); v3_sub(C, P, E);
    to_v3(&f, d, P); v3_sub(D, P, E);
    Big* det = (Big*)malloc(sizeof(Big));
    if (!det) abort();
    det4_lifted(det, A, B, C, D);
    int s = -det->sg;
    free(det);
    free(V);
    return s;
}

int orc_int_in_sphere(const double* a, const double* b, const double* c,
                      const double* d, const double* e) {
    int s = orc_int_in_sphere_raw(a, b, c, d, e);
    if (s != 0) return s;
    /* rank the five argument slots by (coordinates, slot): a stable order */
    const double* p[5] = {a, b, c, d, e};
    int rank[5];
    for (int s1 = 0; s1 < 5; ++s1) {
        int r = 0;
        for (int s2 = 0; s2 < 5; ++s2) {
            if (s2 == s1) continue;
            const double* x = p[s2];
            const double* y = p[s1];
            int less;
            if (x[0] != y[0]) less = x[0] < y[0];
            else if (x[1] != y[1]) less = x[1] < y[1];
            else less = x[2] < y[2];
            int equal = !less && !(y[0] != x[0] ? y[0] < x[0] : (y[1] != x[1] ? y[1] < x[1] : y[2] < x[2]));
            if (less || (equal && s2 < s1)) ++r;
        }
        rank[s1] = r;
    }
    const double* sorted[5];
    for (int s1 = 0; s1 < 5; ++s1) sorted[rank[s1]] = p[s1];
    for (int i = 4; i > 2; --i) {
        if (sorted[i] == e) return -1;
        int o;
        if (sorted[i] == d && (o = orc_int_orient3d(a, b, c, e)) != 0) return o;
        if (sorted[i] == c && (o = orc_int_orient3d(a, b, e, d)) != 0) return o;
        if (sorted[i] == b && (o = orc_int_orient3d(a, e, c, d)) != 0) return o;
        if (sorted[i] == a && (o = orc_int_orient3d(e, b, c, d)) != 0) return o;
    }
    return -1;
}

void orc_int_pred_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                        const int32_t* idx5, int64_t n_q, int8_t* sign_out) {
    (void)n_pts;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const double* a = xyz + 3 * (int64_t)r[0];
        const double* b = xyz + 3 * (int64_t)r[1];
        const double* c = xyz + 3 * (int64_t)r[2];
        const double* d = xyz + 3 * (int64_t)r[3];
        int s;
        if (op[q] == ORC_OP_ORIENT3D) {
            s = orc_int_orient3d(a, b, c, d);
        } else {
            const double* e = xyz + 3 * (int64_t)r[4];
            s = op[q] == ORC_OP_INSPHERE ? orc_int_in_sphere(a, b, c, d, e)
                                         : orc_int_in_sphere_raw(a, b, c, d, e);
        }
        sign_out[q] = (int8_t)s;
    }
}

/* ---- coplanar predicates, method of tests/exact.h:294-378 --------------- */

int orc_int_orient2d(double px, double py, double qx, double qy, double rx, double ry) {
    Frame f = {0, 0};
    frame_add(&f, px); frame_add(&f, py); frame_add(&f, qx);
    frame_add(&f, qy); frame_add(&f, rx); frame_add(&f, ry);
    Big* v = (Big*)malloc(sizeof(Big) * 8);
    if (!v) abort();
    frame_conv(&f, px, v + 0); frame_conv(&f, py, v + 1);
    frame_conv(&f, qx, v + 2); frame_conv(&f, qy, v + 3);
    frame_conv(&f, rx, v + 4); frame_conv(&f, ry, v + 5);
    big_sub(v + 2, v + 2, v + 0); big_sub(v + 3, v + 3, v + 1);  /* q - p */
    big_sub(v + 4, v + 4, v + 0); big_sub(v + 5, v + 5, v + 1);  /* r - p */
    minor2(v + 6, v + 2, v + 5, v + 3, v + 4);
    const int s = v[6].sg;
    free(v);
    return s;
}

static int int_cop_orient3(const double* p, const double* q, const double* r) {
    int o = orc_int_orient2d(p[0], p[1], q[0], q[1], r[0], r[1]);
    if (o) return o;
    o = orc_int_orient2d(p[1], p[2], q[1], q[2], r[1], r[2]);
    if (o) return o;
    return orc_int_orient2d(p[0], p[2], q[0], q[2], r[0], r[2]);
}

static int int_cop_orient4(const double* p, const double* q, const double* r, const double* s) {
    int o = orc_int_orient2d(p[0], p[1], q[0], q[1], r[0], r[1]);
    if (o) return o * orc_int_orient2d(p[0], p[1], q[0], q[1], s[0], s[1]);
    o = orc_int_orient2d(p[1], p[2], q[1], q[2], r[1], r[2]);
    if (o) return o * orc_int_orient2d(p[1], p[2], q[1], q[2], s[1], s[2]);
    o = orc_int_orient2d(p[0], p[2], q[0], q[2], r[0], r[2]);
    return o * orc_int_orient2d(p[0], p[2], q[0], q[2], s[0], s[2]);
}

/* all three projections degenerate, in the order xy, yz, xz of tests/exact.h */
static int int_collinear(const double* p, const double* q, const double* r) {
    return orc_int_orient2d(p[0], p[1], q[0], q[1], r[0], r[1]) == 0 &&
           orc_int_orient2d(p[1], p[2], q[1], q[2], r[1], r[2]) == 0 &&
           orc_int_orient2d(p[0], p[2], q[0], q[2], r[0], r[2]) == 0;
}

/* In-sphere of p against (p0, p1, p2, apex), apex = p0 + (p1-p0) x (p2-p0),
 * all relative to p: -sign det4_lifted(A, B, C, D). */
static int int_in_circle_raw(const double* p0, const double* p1, const double* p2,
                             const double* p) {
    Frame f = {0, 0};
    for (int k = 0; k < 3; ++k) {
        frame_add(&f, p0[k]); frame_add(&f, p1[k]); frame_add(&f, p2[k]); frame_add(&f, p[k]);
    }
    V3* V = (V3*)malloc(sizeof(V3) * 9);
    Big* t = (Big*)malloc(sizeof(Big) * 2);
    if (!V || !t) abort();
    V3 *P = V, *A = V + 1, *B = V + 2, *C = V + 3, *D = V + 4, *T = V + 5, *u = V + 6, *w = V + 7;
    to_v3(&f, p, P);
    to_v3(&f, p0, T); v3_sub(A, T, P);
    to_v3(&f, p1, T); v3_sub(B, T, P);
    to_v3(&f, p2, T); v3_sub(C, T, P);
    v3_sub(u, B, A);
    v3_sub(w, C, A);
    /* n = u x w; D = A + n */
    minor2(&t[0], &u->y, &w->z, &u->z, &w->y); big_add(&D->x, &A->x, &t[0]);
    minor2(&t[0], &u->z, &w->x, &u->x, &w->z); big_add(&D->y, &A->y, &t[0]);
    minor2(&t[0], &u->x, &w->y, &u->y, &w->x); big_add(&D->z, &A->z, &t[0]);
    det4_lifted(&t[1], A, B, C, D);
    const int s = -t[1].sg;
    free(t);
    free(V);
    return s;
}

static int int_in_circle(const double* p0, const double* p1, const double* p2,
                         const double* p) {
    const int bs = int_in_circle_raw(p0, p1, p2, p);
    if (bs != 0) return bs;
    /* rank the four slots by (coordinates, slot): a stable order */
    const double* s[4] = {p0, p1, p2, p};
    const double* sorted[4];
    for (int i = 0; i < 4; ++i) {
        int r = 0;
        for (int j = 0; j < 4; ++j) {
            if (j == i) continue;
            const double* x = s[j];
            const double* y = s[i];
            int c = 0;
            for (int k = 0; k < 3 && c == 0; ++k) c = x[k] < y[k] ? -1 : (x[k] > y[k] ? 1 : 0);
            if (c < 0 || (c == 0 && j < i)) ++r;
        }
        sorted[r] = s[i];
    }
    const int local = int_cop_orient3(p0, p1, p2);
    for (int i = 3; i > 0; --i) {
        if (sorted[i] == p) return -1;
        int o;
        if (sorted[i] == p2 && (o = int_cop_orient3(p0, p1, p)) != 0) return o * local;
        if (sorted[i] == p1 && (o = int_cop_orient3(p0, p, p2)) != 0) return o * local;
        if (sorted[i] == p0 && (o = int_cop_orient3(p, p1, p2)) != 0) return o * local;
    }
    return -local;
}

static int int_equal(const double* p, const double* q) {
    return p[0] == q[0] && p[1] == q[1] && p[2] == q[2];
}

/* Position of q on the line (p, r) from the parameter t of q = p + t (r - p),
 * decided in exact integers along the coordinate of largest extent: a second
 * route to the answer of src/predicates2.cpp:162-180 on collinear input. */
static int int_position(const double* p, const double* q, const double* r) {
    if (int_equal(q, p)) return 1;
    if (int_equal(q, r)) return 3;
    int k = 0;
    for (int j = 1; j < 3; ++j)
        if (fabs(r[j] - p[j]) > fabs(r[k] - p[k])) k = j;
    /* compare (q-p)_k with 0 and with (r-p)_k, in the direction of r-p */
    Frame f = {0, 0};
    frame_add(&f, p[k]); frame_add(&f, q[k]); frame_add(&f, r[k]);
    Big* v = (Big*)malloc(sizeof(Big) * 5);
    if (!v) abort();
    frame_conv(&f, p[k], v); frame_conv(&f, q[k], v + 1); frame_conv(&f, r[k], v + 2);
    big_sub(v + 3, v + 1, v);      /* q - p */
    big_sub(v + 4, v + 2, v);      /* r - p */
    const int dir = v[4].sg;
    big_sub(v + 2, v + 3, v + 4);  /* q - r */
    const int s0 = v[3].sg * dir, s1 = v[2].sg * dir;
    free(v);
    if (s0 < 0) return 0;
    if (s1 > 0) return 4;
    return 2;
}

void orc_int_pred2_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                         const int32_t* idx5, int64_t n_q, int8_t* sign_out) {
    (void)n_pts;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const int np = orc_pred2_points(op[q]);
        const double* p[4];
        for (int k = 0; k < 4; ++k) p[k] = k < np ? xyz + 3 * (int64_t)r[k] : xyz;
        int s;
        switch (op[q]) {
            case ORC_OP2_COPLANAR_ORIENTATION: s = int_cop_orient3(p[0], p[1], p[2]); break;
            case ORC_OP2_COPLANAR_ORIENTATION4: s = int_cop_orient4(p[0], p[1], p[2], p[3]); break;
            case ORC_OP2_COLLINEAR: s = int_collinear(p[0], p[1], p[2]); break;
            case ORC_OP2_EQUAL: s = int_equal(p[0], p[1]); break;
            case ORC_OP2_COLLINEAR_POSITION: s = int_position(p[0], p[1], p[2]); break;
            case ORC_OP2_IN_CIRCLE: s = int_in_circle(p[0], p[1], p[2], p[3]); break;
            case ORC_OP2_IN_CIRCLE_UNPERTURBED: s = int_in_circle_raw(p[0], p[1], p[2], p[3]); break;
            case ORC_OP2_ORIENT2D_XY:
                s = orc_int_orient2d(p[0][0], p[0][1], p[1][0], p[1][1], p[2][0], p[2][1]);
                break;
            case ORC_OP2_INEXACT_ORIENTATION:  /* the exact sign (the op itself is not exact) */
                s = orc_int_orient3d(p[0], p[1], p[2], p[3]);
                break;
            default: s = 0; break;
        }
        sign_out[q] = (int8_t)s;
    }
}
