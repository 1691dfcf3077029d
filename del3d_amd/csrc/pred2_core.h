// pred2_core.h - per-query arithmetic of the coplanar predicates.
//
// The reference's predicates on coplanar points (src/predicates2.cpp:111-259):
// orient2d, coplanar_orientation (three and four points), collinear, equal,
// collinear_position and coplanar_in_circle(_unperturbed) with its symbolic
// perturbation; and the Delaunay driver's plain-double pre-walk orientation
// (src/triangulation.cpp:296-308). As for the 3D predicates, the filters are the reference's,
// operation for operation (compiled with -ffp-contract=off), and whatever
// they leave undecided is decided by exact integer arithmetic on the query's
// own power-of-two grid (the method of tests/exact.h:161-203), whose sign is
// the reference's expansion stage's sign.
//
// The in-circle determinant is not homogeneous: with the rows the reference
// uses (pt, rt, qt, v with v = pq x pr, src/predicates2.cpp:42-101) its terms
// have degree 6, except m012 * |v|^2, of degree 7, which vanishes when the
// query point is coplanar. On a grid of unit 2^q the determinant is
// 2^(6q) * (D6 - E7 * 2^q) for integers D6, E7, and the exact tier takes the
// sign of that combination in the original units, so even a query that
// violates the coplanarity precondition gets the reference's sign.
//
// Every function is __host__ __device__ (host build: csrc/cpu_check.cpp).
#pragma once

#include "pred_core.h"

namespace d3 {

// The exact tiers are large and rarely run: compiled once per translation
// unit and called (inline only for linkage: several sources include this).
#define D3_HD_CALL __host__ __device__ inline __attribute__((noinline))

// src/predicates2.cpp:22-26
constexpr double kOrient2dBound = (3.0 + 16.0 * kEps) * kEps;
constexpr double kInCircleBound = (16.0 + 224.0 * kEps) * kEps;

// Op codes of the coplanar batch (include/del3d_amd.h, DEL3D_OP2_*).
enum : uint8_t {
    kOp2CoplanarOrientation = 0,
    kOp2CoplanarOrientation4 = 1,
    kOp2Collinear = 2,
    kOp2Equal = 3,
    kOp2CollinearPosition = 4,
    kOp2InCircle = 5,
    kOp2InCircleUnperturbed = 6,
    kOp2Orient2dXY = 7,
    kOp2InexactOrientation = 8,
    kOp2Last = 8,
};

// "Not decided without the exact stage" (collinear_position's results use 0-4).
constexpr int kPending2 = 9;

// Points of each op (the rest of a query's indices are ignored).
D3_HD int points_of_op2(uint8_t o) {
    switch (o) {
        case kOp2Equal: return 2;
        case kOp2CoplanarOrientation4:
        case kOp2InCircle:
        case kOp2InCircleUnperturbed:
        case kOp2InexactOrientation: return 4;
        default: return 3;
    }
}

// src/triangulation.cpp:296-308: the driver's pre-walk orientation, the sign
// of det(q-p, r-p, s-p) in plain double precision with the driver's fixed
// association of the products (no filter, no exact stage: near zero its sign
// is the rounded determinant's, which the pre-walk depends on).
D3_HD int inexact_orientation(const P3& p, const P3& q, const P3& r, const P3& s) {
    const double a00 = q.x - p.x, a01 = q.y - p.y, a02 = q.z - p.z;
    const double a10 = r.x - p.x, a11 = r.y - p.y, a12 = r.z - p.z;
    const double a20 = s.x - p.x, a21 = s.y - p.y, a22 = s.z - p.z;
    const double m01 = a00 * a11 - a10 * a01;
    const double m02 = a00 * a21 - a20 * a01;
    const double m12 = a10 * a21 - a20 * a11;
    const double det = m01 * a22 - m02 * a12 + m12 * a02;
    return sgn(det);
}

// ---- filters ---------------------------------------------------------------

// src/predicates2.cpp:182-230: sign if the filter decides, else kPending2.
D3_HD int in_circle_filter(const P3& p0, const P3& p1, const P3& p2, const P3& p, double& det_out) {
    const double ptx = p0.x - p.x, pty = p0.y - p.y, ptz = p0.z - p.z;
    const double qtx = p1.x - p.x, qty = p1.y - p.y, qtz = p1.z - p.z;
    const double rtx = p2.x - p.x, rty = p2.y - p.y, rtz = p2.z - p.z;
    const double pt2 = ptx * ptx + pty * pty + ptz * ptz;
    const double qt2 = qtx * qtx + qty * qty + qtz * qtz;
    const double rt2 = rtx * rtx + rty * rty + rtz * rtz;
    const double pqx = p1.x - p0.x, pqy = p1.y - p0.y, pqz = p1.z - p0.z;
    const double prx = p2.x - p0.x, pry = p2.y - p0.y, prz = p2.z - p0.z;
    const double vx = pqy * prz - pqz * pry;
    const double vy = pqz * prx - pqx * prz;
    const double vz = pqx * pry - pqy * prx;
    const double v2 = vx * vx + vy * vy + vz * vz;
    // rows (pt, rt, qt, v), each lifted; minors in the reference's operand order
    const double a00 = ptx, a01 = pty, a02 = ptz, a03 = pt2;
    const double a10 = rtx, a11 = rty, a12 = rtz, a13 = rt2;
    const double a20 = qtx, a21 = qty, a22 = qtz, a23 = qt2;
    const double a30 = vx, a31 = vy, a32 = vz, a33 = v2;
    const double m01 = a10 * a01 - a00 * a11;
    const double m02 = a20 * a01 - a00 * a21;
    const double m03 = a30 * a01 - a00 * a31;
    const double m12 = a20 * a11 - a10 * a21;
    const double m13 = a30 * a11 - a10 * a31;
    const double m23 = a30 * a21 - a20 * a31;
    const double m012 = m12 * a02 - m02 * a12 + m01 * a22;
    const double m013 = m13 * a02 - m03 * a12 + m01 * a32;
    const double m023 = m23 * a02 - m03 * a22 + m02 * a32;
    const double m123 = m23 * a12 - m13 * a22 + m12 * a32;
    const double det = m123 * a03 - m023 * a13 + m013 * a23 - m012 * a33;
    // magnitude bound: product of the rows' absolute sums, accumulated as the
    // reference does (a zero product restarts at the next row's sum)
    const double rows[4][4] = {{a00, a01, a02, a03}, {a10, a11, a12, a13}, {a20, a21, a22, a23},
                               {a30, a31, a32, a33}};
    double mag = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double row = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) row += fabs(rows[i][j]);
        mag = (mag == 0.0) ? row : mag * row;
    }
    const double errbound = kInCircleBound * mag;
    det_out = det;
    if (det > errbound || -det > errbound) return sgn(det);
    return kPending2;
}

// ---- exact tiers on a caller workspace (runtime limb counts) ----------------

// Grid of a list of doubles.
D3_HD Grid grid_of_values(const double* v, int n) {
    Grid g{0, 0, false};
    for (int i = 0; i < n; ++i) grid_add(g, v[i]);
    return g;
}

// Exact orient2d: sign of (qx-px)(ry-py) - (qy-py)(rx-px). ws: 12*L limbs.
D3_HD_CALL int orient2d_exact(double px, double py, double qx, double qy, double rx, double ry,
                         uint32_t* ws) {
    const double v[6] = {px, py, qx, qy, rx, ry};
    const Grid g = grid_of_values(v, 6);
    const int L = limbs_for(g);
    uint32_t* P = ws;            // px, py
    uint32_t* Q = P + 2 * L;     // qx-px, qy-py
    uint32_t* R = Q + 2 * L;     // rx-px, ry-py
    uint32_t* a = R + 2 * L;     // 2L
    uint32_t* b = a + 2 * L;     // 2L
    uint32_t* tmp = b + 2 * L;   // 2L
    rt_to_limbs(px, g.q, P, L);
    rt_to_limbs(py, g.q, P + L, L);
    rt_to_limbs(qx, g.q, Q, L);
    rt_to_limbs(qy, g.q, Q + L, L);
    rt_to_limbs(rx, g.q, R, L);
    rt_to_limbs(ry, g.q, R + L, L);
    rt_addsub(Q, Q, P, L, 1);
    rt_addsub(Q + L, Q + L, P + L, L, 1);
    rt_addsub(R, R, P, L, 1);
    rt_addsub(R + L, R + L, P + L, L, 1);
    rt_mul(a, Q, L, R + L, L, tmp);
    rt_mul(b, Q + L, L, R, L, tmp);
    rt_addsub(a, a, b, 2 * L, 1);
    return rt_sign(a, 2 * L);
}

// Bit length of |a| for a two's-complement value of n limbs (0 for zero).
D3_HD int rt_bitlen_abs(const uint32_t* a, int n, uint32_t* tmp) {
    rt_cneg(tmp, a, n, a[n - 1] >> 31);
    for (int i = n - 1; i >= 0; --i)
        if (tmp[i]) return 32 * i + 32 - __builtin_clz(tmp[i]);
    return 0;
}

// |a| << s into out (w limbs, zero-filled), a non-negative of n limbs.
D3_HD void rt_shl_into(uint32_t* out, int w, const uint32_t* a, int n, int s) {
    for (int i = 0; i < w; ++i) out[i] = 0;
    const int ws = s / 32, bs = s % 32;
    for (int i = 0; i < n; ++i) {
        const uint64_t x = (uint64_t)a[i] << bs;
        if (i + ws < w) out[i + ws] |= (uint32_t)x;
        if (i + ws + 1 < w) out[i + ws + 1] |= (uint32_t)(x >> 32);
    }
}

// Sign of A - B * 2^q for two's-complement A (na limbs) and B (nb limbs).
// ws: 3 * (na + nb + 3) limbs.
D3_HD int sign_combine(const uint32_t* A, int na, const uint32_t* B, int nb, int q, uint32_t* ws) {
    const int sa = rt_sign(A, na), sb = rt_sign(B, nb);
    if (sb == 0) return sa;
    if (sa == 0) return -sb;
    if (sa != sb) return sa;
    const int w = na + nb + 3;
    uint32_t* t = ws;
    const int la = rt_bitlen_abs(A, na, t);
    const int lb = rt_bitlen_abs(B, nb, t);
    if (la > lb + q) return sa;
    if (la < lb + q) return -sa;
    // equal bit lengths: compare |A| << max(0,-q) with |B| << max(0,q) exactly
    uint32_t* ua = ws + w;
    uint32_t* ub = ws + 2 * w;
    rt_cneg(t, A, na, sa < 0 ? 1u : 0u);
    rt_shl_into(ua, w, t, na, q < 0 ? -q : 0);
    rt_cneg(t, B, nb, sb < 0 ? 1u : 0u);
    rt_shl_into(ub, w, t, nb, q > 0 ? q : 0);
    for (int i = w - 1; i >= 0; --i) {
        if (ua[i] != ub[i]) return ua[i] > ub[i] ? sa : -sa;
    }
    return 0;
}

// Limbs of the in-circle workspace for limb count L: coordinates 12L,
// differences 15L, product and scratch 20L, lifts 6L, normal 6L, |v|^2 4L,
// entries 24L, 2x2 minors 24L, 3x3 minors 24L, D6 8L, E7 10L, and the final
// comparison 3 * (18L + 3).
D3_HD int in_circle_workspace_limbs(int L) { return 207 * L + 16; }
constexpr int kMaxLimbs2 = 67;  // > limbs_for() of any finite input (66)

// Exact coplanar_in_circle_unperturbed determinant sign (src/predicates2.cpp:
// 42-101), in the original units. ws: in_circle_workspace_limbs(L) limbs.
D3_HD_CALL int in_circle_exact(const P3& p0, const P3& p1, const P3& p2, const P3& p, uint32_t* ws) {
    const P3 pts[4] = {p0, p1, p2, p};
    const Grid g = grid_of(pts, 4);
    const int L = limbs_for(g);
    const int q = g.q;
    uint32_t* c = ws;  // 12 coordinates, L each: p0, p1, p2, p
    for (int i = 0; i < 4; ++i) {
        rt_to_limbs(pts[i].x, q, c + (3 * i + 0) * L, L);
        rt_to_limbs(pts[i].y, q, c + (3 * i + 1) * L, L);
        rt_to_limbs(pts[i].z, q, c + (3 * i + 2) * L, L);
    }
#define CO(i, k) (c + (3 * (i) + (k)) * L)
    uint32_t* d = c + 12 * L;  // pt, qt, rt, pq, pr: 15 differences, L each
#define DF(v, k) (d + (3 * (v) + (k)) * L)
    for (int k = 0; k < 3; ++k) {
        rt_addsub(DF(0, k), CO(0, k), CO(3, k), L, 1);  // pt = p0 - p
        rt_addsub(DF(1, k), CO(1, k), CO(3, k), L, 1);  // qt = p1 - p
        rt_addsub(DF(2, k), CO(2, k), CO(3, k), L, 1);  // rt = p2 - p
        rt_addsub(DF(3, k), CO(1, k), CO(0, k), L, 1);  // pq = p1 - p0
        rt_addsub(DF(4, k), CO(2, k), CO(0, k), L, 1);  // pr = p2 - p0
    }
    uint32_t* prod = d + 15 * L;    // 10L (widest product: 6L x 4L)
    uint32_t* tmp = prod + 10 * L;  // 10L
    uint32_t* lift = tmp + 10 * L;  // pt2, qt2, rt2: 2L each
    for (int v = 0; v < 3; ++v) {
        uint32_t* lv = lift + 2 * v * L;
        for (int i = 0; i < 2 * L; ++i) lv[i] = 0;
        for (int k = 0; k < 3; ++k) rt_mac(lv, DF(v, k), L, DF(v, k), L, 0, prod, tmp);
    }
    uint32_t* vv = lift + 6 * L;  // vx, vy, vz: 2L each
    // vx = pqy*prz - pqz*pry ; vy = pqz*prx - pqx*prz ; vz = pqx*pry - pqy*prx
    const int cr[3][4] = {{1, 2, 2, 1}, {2, 0, 0, 2}, {0, 1, 1, 0}};
    for (int k = 0; k < 3; ++k) {
        uint32_t* vk = vv + 2 * k * L;
        rt_mul(vk, DF(3, cr[k][0]), L, DF(4, cr[k][1]), L, tmp);
        rt_mac(vk, DF(3, cr[k][2]), L, DF(4, cr[k][3]), L, 1, prod, tmp);
    }
    uint32_t* v2 = vv + 6 * L;  // 4L
    for (int i = 0; i < 4 * L; ++i) v2[i] = 0;
    for (int k = 0; k < 3; ++k) rt_mac(v2, vv + 2 * k * L, 2 * L, vv + 2 * k * L, 2 * L, 0, prod, tmp);
    // Every entry widened to 2L limbs: rows pt, rt, qt, v (columns x, y, z);
    // the lifts are kept aside (column w).
    uint32_t* e = v2 + 4 * L;  // 12 entries x 2L
#define ENT(r, k) (e + (3 * (r) + (k)) * 2 * L)
    const int rowv[3] = {0, 2, 1};  // rows 0..2 are pt, rt, qt (r before q)
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            // sign-extend the L-limb difference to 2L limbs
            const uint32_t* s = DF(rowv[r], k);
            uint32_t* t = ENT(r, k);
            const uint32_t ext = (s[L - 1] >> 31) ? 0xFFFFFFFFu : 0u;
            for (int i = 0; i < 2 * L; ++i) t[i] = i < L ? s[i] : ext;
        }
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 2 * L; ++i) ENT(3, k)[i] = vv[2 * k * L + i];
    // 2x2 minors of columns 0, 1 over row pairs: m_ij = a_j0*a_i1 - a_i0*a_j1 (4L)
    uint32_t* mn = e + 24 * L;  // m01 m02 m03 m12 m13 m23, 4L each
    const int pr2[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
    for (int m = 0; m < 6; ++m) {
        const int i = pr2[m][0], j = pr2[m][1];
        uint32_t* o = mn + 4 * m * L;
        rt_mul(o, ENT(j, 0), 2 * L, ENT(i, 1), 2 * L, tmp);
        rt_mac(o, ENT(i, 0), 2 * L, ENT(j, 1), 2 * L, 1, prod, tmp);
    }
#define MN(m) (mn + 4 * (m) * L)
    // 3x3 minors (6L): m012 = m12*a02 - m02*a12 + m01*a22 ; m013 = m13*a02 - m03*a12 + m01*a32
    //                  m023 = m23*a02 - m03*a22 + m02*a32 ; m123 = m23*a12 - m13*a22 + m12*a32
    uint32_t* m3 = mn + 24 * L;  // 4 x 6L
    const int t3[4][3][2] = {{{3, 0}, {1, 1}, {0, 2}},   // (minor, row of column-2 entry)
                             {{4, 0}, {2, 1}, {0, 3}},
                             {{5, 0}, {2, 2}, {1, 3}},
                             {{5, 1}, {4, 2}, {3, 3}}};
    for (int k = 0; k < 4; ++k) {
        uint32_t* o = m3 + 6 * k * L;
        for (int i = 0; i < 6 * L; ++i) o[i] = 0;
        for (int t = 0; t < 3; ++t)
            rt_mac(o, MN(t3[k][t][0]), 4 * L, ENT(t3[k][t][1], 2), 2 * L, t == 1 ? 1 : 0, prod, tmp);
    }
    // D6 = m123*pt2 - m023*rt2 + m013*qt2 (8L), E7 = m012 * v2 (10L)
    uint32_t* D6 = m3 + 24 * L;   // 8L
    uint32_t* E7 = D6 + 8 * L;    // 10L
    for (int i = 0; i < 8 * L; ++i) D6[i] = 0;
    rt_mac(D6, m3 + 18 * L, 6 * L, lift + 0 * L, 2 * L, 0, prod, tmp);  // m123 * pt2
    rt_mac(D6, m3 + 12 * L, 6 * L, lift + 4 * L, 2 * L, 1, prod, tmp);  // - m023 * rt2
    rt_mac(D6, m3 + 6 * L, 6 * L, lift + 2 * L, 2 * L, 0, prod, tmp);   // + m013 * qt2
    rt_mul(E7, m3, 6 * L, v2, 4 * L, tmp);                              // m012 * v2
#undef MN
#undef ENT
#undef DF
#undef CO
    // det = 2^(6q) * (D6 - E7 * 2^q)
    return sign_combine(D6, 8 * L, E7, 10 * L, q, E7 + 10 * L);
}

// ---- T1: certified FP64 evaluation on a small power-of-two grid --------------
//
// As for the 3D predicates: with M the largest |difference| the filter forms
// from the query point and q = ilogb(M) - k, if every coordinate is an integer
// multiple of 2^q below 2^(52+q) in magnitude, every difference is exact and
// below 2^(k+1) units of 2^q (the in-circle's pq, pr below 2^(k+2)), and every
// product and partial sum of the filter's expression is an integer below 2^53
// in units of 2^(deg*q): the FP64 determinant is exact, so its sign (zero
// included) is the exact stage's sign.
//   orient2d (degree 2): k = 25, |det| < 2 * 2^52 = 2^53; q >= -537.
//   in-circle (degree 7 for the |v|^2 term): k = 5, a = 2^6 bounds the
//     differences, and the terms sum to below 432 a^6 + 1152 a^7 < 2^52.2;
//     7q >= -1074 needs q >= -153, and 53 + 7q < 1024 needs q <= 138 (e <= 140).
template <int N>
D3_HD bool t1_certify_values(const double (&v)[N], double maxd, int k, int qmin, int emax) {
    const int e = exponent_of(maxd);
    const int q = e - k;
    const bool in_range = (q >= qmin) & (e <= emax);
    const double inv = pow2i(in_range ? -q : 0);
    bool ok = in_range;
#pragma unroll
    for (int i = 0; i < N; ++i) ok = ok & on_grid(v[i], inv);
    return (maxd == 0.0) | ok;
}

// orient2d without the exact stage: the sign when the filter or the
// certificate decides, else kPending2 (src/predicates2.cpp:111-118).
D3_HD int orient2d_fast(double px, double py, double qx, double qy, double rx, double ry) {
    const double qpx = qx - px, qpy = qy - py;
    const double rpx = rx - px, rpy = ry - py;
    const double a = qpx * rpy, b = qpy * rpx;
    const double det = a - b;
    const double permanent = fabs(a) + fabs(b);
    const double errbound = kOrient2dBound * permanent;
    if ((det > errbound) | (-det > errbound)) return sgn(det);
    const double maxd = fmax(fmax(fabs(qpx), fabs(qpy)), fmax(fabs(rpx), fabs(rpy)));
    const double v[6] = {px, py, qx, qy, rx, ry};
    return t1_certify_values(v, maxd, 25, -537, 180) ? sgn(det) : kPending2;
}

// The in-circle determinant's sign without the exact stage, else kPending2.
D3_HD int in_circle_fast(const P3& p0, const P3& p1, const P3& p2, const P3& p) {
    double det;
    const int f = in_circle_filter(p0, p1, p2, p, det);
    if (f != kPending2) return f;
    const P3 pts[4] = {p0, p1, p2, p};
    const double maxd = max_diff(pts, 4);
    const double v[12] = {p0.x, p0.y, p0.z, p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p.x, p.y, p.z};
    return t1_certify_values(v, maxd, 5, -153, 140) ? sgn(det) : kPending2;
}

// ---- the predicates -----------------------------------------------------------
//
// Each predicate is written once over an orient2d evaluator O and run in two
// modes: kFast (filters and certificates only; any undecided step makes the
// whole query kPending2, for the exact kernel) and full (exact tiers behind them).

template <bool kFast>
struct Orient2 {
    uint32_t* ws;
    D3_HD int operator()(double px, double py, double qx, double qy, double rx, double ry) const {
        const int f = orient2d_fast(px, py, qx, qy, rx, ry);
        if (kFast || f != kPending2) return f;
        return orient2d_exact(px, py, qx, qy, rx, ry, ws);
    }
};

// src/predicates2.cpp:122-131
template <class O>
D3_HD int coplanar_orientation3(const P3& p, const P3& q, const P3& r, const O& o2) {
    int o = o2(p.x, p.y, q.x, q.y, r.x, r.y);
    if (o != 0) return o;
    o = o2(p.y, p.z, q.y, q.z, r.y, r.z);
    if (o != 0) return o;
    return o2(p.x, p.z, q.x, q.z, r.x, r.z);
}

// src/predicates2.cpp:133-146
template <class O>
D3_HD int coplanar_orientation4(const P3& p, const P3& q, const P3& r, const P3& s, const O& o2) {
    int o = o2(p.x, p.y, q.x, q.y, r.x, r.y);
    if (o == kPending2) return kPending2;
    if (o != 0) {
        const int t = o2(p.x, p.y, q.x, q.y, s.x, s.y);
        return t == kPending2 ? kPending2 : o * t;
    }
    o = o2(p.y, p.z, q.y, q.z, r.y, r.z);
    if (o == kPending2) return kPending2;
    if (o != 0) {
        const int t = o2(p.y, p.z, q.y, q.z, s.y, s.z);
        return t == kPending2 ? kPending2 : o * t;
    }
    o = o2(p.x, p.z, q.x, q.z, r.x, r.z);
    if (o == kPending2) return kPending2;
    const int t = o2(p.x, p.z, q.x, q.z, s.x, s.z);
    return t == kPending2 ? kPending2 : o * t;
}

// src/predicates2.cpp:148-156 (1 collinear, 0 not, kPending2 undecided)
template <class O>
D3_HD int collinear3(const P3& p, const P3& q, const P3& r, const O& o2) {
    int o = o2(r.x, r.y, p.x, p.y, q.x, q.y);
    if (o != 0) return o == kPending2 ? kPending2 : 0;
    o = o2(r.x, r.z, p.x, p.z, q.x, q.z);
    if (o != 0) return o == kPending2 ? kPending2 : 0;
    o = o2(r.y, r.z, p.y, p.z, q.y, q.z);
    return o == kPending2 ? kPending2 : (o == 0 ? 1 : 0);
}

// src/predicates2.cpp:158-160
D3_HD bool equal3(const P3& p, const P3& q) { return p.x == q.x && p.y == q.y && p.z == q.z; }

// src/predicates2.cpp:162-180: 0 before, 1 source, 2 middle, 3 target, 4 after
D3_HD int collinear_position3(const P3& p, const P3& q, const P3& r) {
    if (equal3(q, p)) return 1;
    if (equal3(q, r)) return 3;
    // the first coordinate in which p and r differ (x, else y, else z),
    // picked with bit masks: a select between the fields would let the
    // compiler select their addresses instead, which keeps the points in
    // scratch memory
    const bool kx = p.x != r.x, ky = !kx && p.y != r.y;
    const uint64_t mx = 0ull - (uint64_t)kx, my = 0ull - (uint64_t)ky, mz = ~(mx | my);
    auto pick = [&](const P3& v) {
        return __builtin_bit_cast(double, (__builtin_bit_cast(uint64_t, v.x) & mx) |
                                              (__builtin_bit_cast(uint64_t, v.y) & my) |
                                              (__builtin_bit_cast(uint64_t, v.z) & mz));
    };
    const double a = pick(p), b = pick(q), c = pick(r);
    if (a < c) {
        if (b < a) return 0;
        if (b > c) return 4;
        return 2;
    }
    if (b > a) return 0;
    if (b < c) return 4;
    return 2;
}

// src/predicates2.cpp:236-259: the cocircular perturbation. Slots 0..3 are
// (p0, p1, p2, p); id[] gives their identities (equal index = same pointer).
template <class O>
D3_HD int in_circle_perturb(const P3 pt[4], const int32_t id[4], const O& o2) {
    // stable lexicographic order of the four slots (std::sort on four
    // elements is an insertion sort, so equal points keep argument order):
    // slot s ends at rank = points strictly below it + equal points before it.
    // Ranks and the selects below use only constant indices, so the points
    // stay in registers.
    int rank[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        int r = 0;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            if (t == s) continue;
            const bool lt = lex_less(pt[t], pt[s]);
            const bool eq = !lt && !lex_less(pt[s], pt[t]);
            r += (lt || (eq && t < s)) ? 1 : 0;
        }
        rank[s] = r;
    }
    const int local = coplanar_orientation3(pt[0], pt[1], pt[2], o2);
    if (local == kPending2) return kPending2;
#pragma unroll
    for (int i = 3; i > 0; --i) {
        int32_t v = id[0];
        v = rank[1] == i ? id[1] : v;
        v = rank[2] == i ? id[2] : v;
        v = rank[3] == i ? id[3] : v;
        if (v == id[3]) return -1;
        int o;
        if (v == id[2] && (o = coplanar_orientation3(pt[0], pt[1], pt[3], o2)) != 0)
            return o == kPending2 ? kPending2 : o * local;
        if (v == id[1] && (o = coplanar_orientation3(pt[0], pt[3], pt[2], o2)) != 0)
            return o == kPending2 ? kPending2 : o * local;
        if (v == id[0] && (o = coplanar_orientation3(pt[3], pt[1], pt[2], o2)) != 0)
            return o == kPending2 ? kPending2 : o * local;
    }
    return -local;
}

// One query, every op. pt: the op's points in argument order; id: their
// identities. kFast: result or kPending2 (undecided by filters and certificates);
// full: the result (ws: in_circle_workspace_limbs(kMaxLimbs2) limbs cover
// every finite input).
template <bool kFast>
D3_HD int pred2_run(uint8_t op, const P3 pt[4], const int32_t id[4], uint32_t* ws) {
    const Orient2<kFast> o2{ws};
    switch (op) {
        case kOp2CoplanarOrientation: return coplanar_orientation3(pt[0], pt[1], pt[2], o2);
        case kOp2CoplanarOrientation4: return coplanar_orientation4(pt[0], pt[1], pt[2], pt[3], o2);
        case kOp2Collinear: return collinear3(pt[0], pt[1], pt[2], o2);
        case kOp2Equal: return equal3(pt[0], pt[1]) ? 1 : 0;
        case kOp2CollinearPosition: return collinear_position3(pt[0], pt[1], pt[2]);
        case kOp2Orient2dXY: return o2(pt[0].x, pt[0].y, pt[1].x, pt[1].y, pt[2].x, pt[2].y);
        case kOp2InexactOrientation: return inexact_orientation(pt[0], pt[1], pt[2], pt[3]);
        case kOp2InCircle:
        case kOp2InCircleUnperturbed: {
            int s = in_circle_fast(pt[0], pt[1], pt[2], pt[3]);
            if (s == kPending2) {
                if (kFast) return kPending2;
                s = in_circle_exact(pt[0], pt[1], pt[2], pt[3], ws);
            }
            if (s != 0 || op == kOp2InCircleUnperturbed) return s;
            return in_circle_perturb(pt, id, o2);
        }
        default: return 0;
    }
}

D3_HD_CALL int pred2_eval(uint8_t op, const P3 pt[4], const int32_t id[4], uint32_t* ws) {
    return pred2_run<false>(op, pt, id, ws);
}

D3_HD int pred2_fast(uint8_t op, const P3 pt[4], const int32_t id[4]) {
    return pred2_run<true>(op, pt, id, nullptr);
}

}  // namespace d3
