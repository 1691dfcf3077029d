// circ_core.h - per-cell arithmetic of the circumcentres.
//
// The reference's two circumcentres of a tetrahedron (src/triangulation.cpp:
// 805-930, include/del3d/delaunay.h:29-49):
//
//   circumcenter       numerator and denominator of the closed form in exact
//                      expansion arithmetic, one rounding per coordinate at
//                      the end (p0 + n.to_double() / d.to_double());
//   circumcenter_lazy  the same formula (Cramer form, p as the origin) in
//                      interval arithmetic under upward rounding, accepting
//                      per coordinate the interval's bound (a point interval)
//                      or its midpoint (relative radius below 1e-5), else the
//                      exact circumcentre's coordinate.
//
// Bit-compatibility:
//  * The exact form's result depends on the expansions' component lists (the
//    final to_double sums them smallest first), so the expansion operations
//    here are the reference's algorithms step for step (src/expansion.h:
//    33-178): the same two_sum chains, the same zero elimination, the same
//    order of scaling and accumulation. The round-off terms are exact, so
//    two_product's error is computed with one fused multiply-add; it is the
//    same double Dekker's splitting produces.
//  * The interval pass needs round-toward-+infinity for every operation
//    (src/interval.h). The GPU keeps round-to-nearest and derives each
//    upward-rounded result from the nearest one and its exact error (a
//    two_sum, or a fused multiply-add for products and quotients): RU(x) is
//    the next double above RN(x) exactly when the error is positive. This is
//    exact for finite results whose error terms do not underflow (operands
//    above about 2^-960 in magnitude), the range the reference's own
//    expansion arithmetic needs as well.
//
// Every function is __host__ __device__ (host build: csrc/cpu_check.cpp).
#pragma once

#include "pred_core.h"

namespace d3 {

// ---- upward rounding from round-to-nearest -----------------------------------

D3_HD double next_up(double x) {
    const uint64_t b = __builtin_bit_cast(uint64_t, x);
    if (x == 0.0) return __builtin_bit_cast(double, (uint64_t)1);  // +2^-1074
    return __builtin_bit_cast(double, (b >> 63) ? b - 1 : b + 1);
}

constexpr double kDblMax = 1.7976931348623157e308;

// A negative result that overflows rounds upward to -DBL_MAX, not -inf.
D3_HD double ru_overflow_fix(double r, double a, double b) {
    return (r == -__builtin_inf() && fabs(a) <= kDblMax && fabs(b) <= kDblMax) ? -kDblMax : r;
}

D3_HD double ru_add(double a, double b) {
    const double s = a + b;
    const double bv = s - a;
    const double av = s - bv;
    const double e = (a - av) + (b - bv);
    if (fabs(s) > kDblMax) return ru_overflow_fix(s, a, b);
    return e > 0.0 ? next_up(s) : s;
}

D3_HD double ru_mul(double a, double b) {
    const double p = a * b;
    if (fabs(p) > kDblMax) return ru_overflow_fix(p, a, b);
    const double e = __builtin_fma(a, b, -p);
    return e > 0.0 ? next_up(p) : p;
}

D3_HD double ru_div(double a, double b) {
    const double q = a / b;
    if (fabs(q) > kDblMax) return b == 0.0 ? q : ru_overflow_fix(q, a, b);
    if (b == 0.0 || !(fabs(q) <= kDblMax)) return q;
    const double r = __builtin_fma(-q, b, a);  // a - q*b, exact
    // the true quotient exceeds q when the remainder has the divisor's sign
    return (r != 0.0 && ((r > 0.0) == (b > 0.0))) ? next_up(q) : q;
}

// ---- interval arithmetic (src/interval.h:52-186) -----------------------------
// Stored as the reference stores it, the lower bound negated, so that each
// bound is one upward-rounded operation; the case analysis of the product,
// quotient and square is the reference's (it decides which corner products
// are formed, and so the bounds' bits and zero signs).

struct Ival {
    double ni;  // -inf
    double s;   // sup
    D3_HD double lo() const { return -ni; }
    D3_HD double hi() const { return s; }
};

D3_HD Ival ival(double d) { return Ival{-d, d}; }
D3_HD Ival ival2(double i, double s) { return Ival{-i, s}; }

D3_HD Ival operator+(const Ival& a, const Ival& b) {
    return ival2(-ru_add(-a.lo(), -b.lo()), ru_add(a.hi(), b.hi()));
}

D3_HD Ival operator-(const Ival& a, const Ival& b) {
    return ival2(-ru_add(-a.lo(), b.hi()), ru_add(a.hi(), -b.lo()));
}

D3_HD Ival operator*(const Ival& a, const Ival& b) {
    if (a.lo() >= 0.0) {
        double aa = a.lo(), bb = a.hi();
        if (bb <= 0.0) return ival(0.0);
        if (b.lo() < 0.0) {
            aa = bb;
            if (b.hi() < 0.0) bb = a.lo();
        }
        const double r = (b.hi() == 0) ? 0.0 : ru_mul(bb, b.hi());
        return ival2(-ru_mul(aa, -b.lo()), r);
    }
    if (a.hi() <= 0.0) {
        double aa = a.hi(), bb = a.lo();
        if (b.lo() < 0.0) {
            aa = bb;
            if (b.hi() <= 0.0) bb = a.hi();
        } else if (b.hi() <= 0) {
            return ival(0.0);
        }
        return ival2(-ru_mul(-bb, b.hi()), ru_mul(-aa, -b.lo()));
    }
    if (b.lo() >= 0.0) {
        if (b.hi() <= 0.0) return ival(0.0);
        return ival2(-ru_mul(-a.lo(), b.hi()), ru_mul(a.hi(), b.hi()));
    }
    if (b.hi() <= 0.0) return ival2(-ru_mul(a.hi(), -b.lo()), ru_mul(-a.lo(), -b.lo()));
    const double t1 = ru_mul(-a.lo(), b.hi());
    const double t2 = ru_mul(a.hi(), -b.lo());
    const double t3 = ru_mul(-a.lo(), -b.lo());
    const double t4 = ru_mul(a.hi(), b.hi());
    return ival2(-(t1 < t2 ? t2 : t1), t3 < t4 ? t4 : t3);
}

D3_HD Ival operator/(const Ival& a, const Ival& b) {
    if (b.lo() > 0.0) {
        double aa = b.hi(), bb = b.lo();
        if (a.lo() < 0.0) {
            aa = bb;
            if (a.hi() < 0.0) bb = b.hi();
        }
        return ival2(-ru_div(-a.lo(), aa), ru_div(a.hi(), bb));
    }
    if (b.hi() < 0.0) {
        double aa = b.hi(), bb = b.lo();
        if (a.lo() < 0.0) {
            bb = aa;
            if (a.hi() < 0.0) aa = b.lo();
        }
        return ival2(-ru_div(a.hi(), -aa), ru_div(a.lo(), bb));
    }
    return ival2(-__builtin_inf(), __builtin_inf());
}

D3_HD Ival square(const Ival& d) {
    if (d.lo() >= 0.0) return ival2(-ru_mul(-d.lo(), d.lo()), ru_mul(d.hi(), d.hi()));
    if (d.hi() <= 0.0) return ival2(-ru_mul(d.hi(), -d.hi()), ru_mul(-d.lo(), -d.lo()));
    const double m = (-d.lo() < d.hi()) ? d.hi() : -d.lo();
    return ival2(0.0, ru_mul(m, m));
}

// src/interval.h:158-174, all under upward rounding
D3_HD double ival_mid(const Ival& v) { return ru_mul(ru_add(v.hi(), v.lo()), 0.5); }

D3_HD bool ival_accurate(const Ival& v, double prec) {
    const double a = fabs(v.lo()), b = fabs(v.hi());
    const double mag = a < b ? b : a;
    const double radius = ru_div(ru_add(v.hi(), -v.lo()), 2.0);
    return mag == 0 || radius < ru_mul(prec, mag);
}

// src/triangulation.cpp:861-869: the 3x3 determinant with its fixed association
D3_HD Ival det3i(const Ival& a00, const Ival& a01, const Ival& a02, const Ival& a10, const Ival& a11,
                 const Ival& a12, const Ival& a20, const Ival& a21, const Ival& a22) {
    const Ival m01 = a00 * a11 - a10 * a01;
    const Ival m02 = a00 * a21 - a20 * a01;
    const Ival m12 = a10 * a21 - a20 * a11;
    return m01 * a22 - m02 * a12 + m12 * a02;
}

// The interval pass of circumcenter_lazy (src/triangulation.cpp:874-920).
// Writes the accepted coordinates to out[] and returns a 3-bit mask of the
// coordinates that need the exact circumcentre.
D3_HD int circ_lazy_pass(const P3& p, const P3& q, const P3& r, const P3& s, double out[3]) {
    const Ival px = ival(p.x), py = ival(p.y), pz = ival(p.z);
    const Ival qpx = ival(q.x) - px, qpy = ival(q.y) - py, qpz = ival(q.z) - pz;
    const Ival qp2 = square(qpx) + square(qpy) + square(qpz);
    const Ival rpx = ival(r.x) - px, rpy = ival(r.y) - py, rpz = ival(r.z) - pz;
    const Ival rp2 = square(rpx) + square(rpy) + square(rpz);
    const Ival spx = ival(s.x) - px, spy = ival(s.y) - py, spz = ival(s.z) - pz;
    const Ival sp2 = square(spx) + square(spy) + square(spz);
    const Ival num_x = det3i(qpy, qpz, qp2, rpy, rpz, rp2, spy, spz, sp2);
    const Ival num_y = det3i(qpx, qpz, qp2, rpx, rpz, rp2, spx, spz, sp2);
    const Ival num_z = det3i(qpx, qpy, qp2, rpx, rpy, rp2, spx, spy, sp2);
    const Ival den = det3i(qpx, qpy, qpz, rpx, rpy, rpz, spx, spy, spz);
    const Ival inv = ival(1.0) / (ival(2.0) * den);
    const Ival c[3] = {px + num_x * inv, py - num_y * inv, pz + num_z * inv};
    const double kPrec = 0.00001;
    int need = 0;
    for (int i = 0; i < 3; ++i) {
        if (c[i].hi() == c[i].lo()) out[i] = c[i].lo();
        else if (ival_accurate(c[i], kPrec)) out[i] = ival_mid(c[i]);
        else need |= 1 << i;
    }
    return need;
}

// ---- expansions on a caller workspace (src/expansion.h:33-178) ---------------
// An expansion is (pointer, length); the workspace is strided so that lanes
// working on different cells touch consecutive addresses for the same element
// (element i of a lane's buffer is at p[i * stride]).

struct Buf {
    double* p;
    int64_t stride;
    D3_HD double& operator[](int i) const { return p[(int64_t)i * stride]; }
    D3_HD Buf at(int off) const { return Buf{p + (int64_t)off * stride, stride}; }
};

D3_HD void two_sum(double a, double b, double& x, double& err) {
    x = a + b;
    const double bv = x - a;
    const double av = x - bv;
    err = (a - av) + (b - bv);
}

// from_diff: the exact a - b (at most 2 components)
D3_HD int x_diff(double a, double b, Buf out) {
    double x, e;
    two_sum(a, -b, x, e);
    int n = 0;
    if (e != 0.0) out[n++] = e;
    if (x != 0.0) out[n++] = x;
    return n;
}

// grow: out = x + b (out must not overlap x)
D3_HD int x_grow(Buf x, int n, double b, Buf out) {
    int m = 0;
    double q = b;
    for (int i = 0; i < n; ++i) {
        double sum, err;
        two_sum(q, x[i], sum, err);
        if (err != 0.0) out[m++] = err;
        q = sum;
    }
    if (q != 0.0) out[m++] = q;
    return m;
}

// x + y (or x - y): y's components grown into x one at a time, ping-ponging
// between t0 and t1 (each of capacity n + m); returns the buffer holding the
// result in *res. x may be t0 or t1 only if it is *res of a previous call and
// the other buffer is free: callers pass distinct buffers.
D3_HD int x_add(Buf x, int n, Buf y, int m, bool negate_y, Buf t0, Buf t1, Buf* res) {
    Buf cur = x;
    int len = n;
    bool use0 = true;
    for (int i = 0; i < m; ++i) {
        const double b = negate_y ? -y[i] : y[i];
        if (b == 0.0) continue;
        Buf out = use0 ? t0 : t1;
        len = x_grow(cur, len, b, out);
        cur = out;
        use0 = !use0;
    }
    *res = cur;
    return len;
}

// scaled: out = x * b (at most 2n components)
D3_HD int x_scale(Buf x, int n, double b, Buf out) {
    if (b == 0.0 || n == 0) return 0;
    int m = 0;
    double q = x[0] * b;
    double hh = __builtin_fma(x[0], b, -q);
    if (hh != 0.0) out[m++] = hh;
    for (int i = 1; i < n; ++i) {
        const double pi = x[i] * b;
        const double pe = __builtin_fma(x[i], b, -pi);
        double s, e;
        two_sum(q, pe, s, e);
        if (e != 0.0) out[m++] = e;
        double t, te;
        two_sum(pi, s, t, te);
        if (te != 0.0) out[m++] = te;
        q = t;
    }
    if (q != 0.0) out[m++] = q;
    return m;
}

// x * y: acc = acc + scaled(x, y_i) for each component of y. sc holds 2n, a0
// and a1 hold 2nm each. Returns the buffer holding the product in *res.
D3_HD int x_mul(Buf x, int n, Buf y, int m, Buf sc, Buf a0, Buf a1, Buf* res) {
    Buf acc = a0;
    int len = 0;
    bool in0 = true;
    for (int i = 0; i < m; ++i) {
        const int k = x_scale(x, n, y[i], sc);
        // acc + sc, grown component by component, ping-ponging a0/a1
        for (int j = 0; j < k; ++j) {
            Buf out = in0 ? a1 : a0;
            len = x_grow(acc, len, sc[j], out);
            acc = out;
            in0 = !in0;
        }
    }
    *res = acc;
    return len;
}

D3_HD void x_copy(Buf from, int n, Buf to) {
    for (int i = 0; i < n; ++i) to[i] = from[i];
}

D3_HD double x_to_double(Buf x, int n) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += x[i];
    return s;
}

// Workspace layout of the exact circumcentre (doubles per lane). The first
// three regions have fixed worst-case sizes; the product, term and sum buffers
// are sized by a layout: kCircFull holds the worst case of every step
// (2 n m components for a product, n + m for a sum), kCircSmall a fraction of
// it that typical cells fit, with every step's bound checked against it.
constexpr int kCircDiff = 0;                   // 9 differences x 2
constexpr int kCircCross = kCircDiff + 18;     // 9 cross-product coordinates x 16
constexpr int kCircSq = kCircCross + 144;      // |a|^2, |b|^2, |c|^2 x 24
constexpr int kCircFixed = kCircSq + 72;

struct CircLayout {
    int cap_sc, cap_m, cap_a;  // scaled buffer, product/term buffers, sum buffers
    D3_HD constexpr int sc() const { return kCircFixed; }
    D3_HD constexpr int m0() const { return sc() + cap_sc; }
    D3_HD constexpr int m1() const { return m0() + cap_m; }
    D3_HD constexpr int t() const { return m1() + cap_m; }
    D3_HD constexpr int a0() const { return t() + cap_m; }
    D3_HD constexpr int a1() const { return a0() + cap_a; }
    D3_HD constexpr int n() const { return a1() + cap_a; }
    D3_HD constexpr int total() const { return n() + cap_a; }
};
// worst case: scaled 2 x 192 (the denominator's x 2), products 2 x 24 x 16,
// sums 3 x 768
constexpr CircLayout kCircFull{384, 768, 2304};
constexpr CircLayout kCircSmall{96, 128, 256};

// a*b - c*d of two-component differences (at most 16 components) into out
D3_HD int x_cross_term(Buf a, int na, Buf b, int nb, Buf c, int nc, Buf d, int nd, Buf ws,
                       const CircLayout& L, Buf out) {
    Buf r1, r2;
    const int l1 = x_mul(a, na, b, nb, ws.at(L.sc()), ws.at(L.m0()), ws.at(L.m1()), &r1);
    x_copy(r1, l1, ws.at(L.t()));
    const int l2 = x_mul(c, nc, d, nd, ws.at(L.sc()), ws.at(L.m0()), ws.at(L.m1()), &r2);
    Buf res;
    const int l = x_add(ws.at(L.t()), l1, r2, l2, true, ws.at(L.a0()), ws.at(L.a1()), &res);
    x_copy(res, l, out);
    return l;
}

// sum = sum + x * y over the layout's buffers (sum lives in N, length *ln;
// first: the product becomes the sum). False if a bound exceeds the layout.
D3_HD bool x_accumulate_product(Buf x, int nx, Buf y, int ny, bool first, Buf ws, const CircLayout& L,
                                int* ln) {
    if (2 * nx > L.cap_sc || 2 * nx * ny > L.cap_m) return false;
    Buf r;
    const int l = x_mul(x, nx, y, ny, ws.at(L.sc()), ws.at(L.m0()), ws.at(L.m1()), &r);
    if (first) {
        x_copy(r, l, ws.at(L.n()));
        *ln = l;
        return true;
    }
    if (*ln + l > L.cap_a) return false;
    x_copy(r, l, ws.at(L.t()));
    Buf s;
    *ln = x_add(ws.at(L.n()), *ln, ws.at(L.t()), l, false, ws.at(L.a0()), ws.at(L.a1()), &s);
    x_copy(s, *ln, ws.at(L.n()));
    return true;
}

// The exact circumcentre (src/triangulation.cpp:805-852). ws: L.total()
// doubles. *overflow is set (and p0 returned) if a step's component bound
// exceeds the layout; with kCircFull it never is.
D3_HD P3 circ_exact(const P3& p0, const P3& p1, const P3& p2, const P3& p3, Buf ws, const CircLayout& L,
                    bool* overflow) {
    *overflow = false;
    // a = p1-p0, b = p2-p0, c = p3-p0: diff k (0..8) = a.x a.y a.z b.x .. c.z
    const double src[9] = {p1.x, p1.y, p1.z, p2.x, p2.y, p2.z, p3.x, p3.y, p3.z};
    const double org[3] = {p0.x, p0.y, p0.z};
    int nd[9];
    for (int k = 0; k < 9; ++k) nd[k] = x_diff(src[k], org[k % 3], ws.at(kCircDiff + 2 * k));
#define DIFF(k) ws.at(kCircDiff + 2 * (k)), nd[k]
    // b x c, c x a, a x b (vectors a 0, b 1, c 2; coordinates x 0, y 1, z 2):
    // (u x v)_k = u_{k+1} v_{k+2} - u_{k+2} v_{k+1}
    const int pairs[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    int nc[9];
    for (int w = 0; w < 3; ++w) {
        const int u = pairs[w][0], v = pairs[w][1];
        for (int k = 0; k < 3; ++k) {
            const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
            nc[3 * w + k] = x_cross_term(DIFF(3 * u + k1), DIFF(3 * v + k2), DIFF(3 * u + k2),
                                         DIFF(3 * v + k1), ws, L, ws.at(kCircCross + 16 * (3 * w + k)));
        }
    }
#define CROSS(w, k) ws.at(kCircCross + 16 * (3 * (w) + (k))), nc[3 * (w) + (k)]
    // |a|^2, |b|^2, |c|^2 = x*x + y*y + z*z, summed left to right (at most 24)
    int nsq[3];
    for (int v = 0; v < 3; ++v) {
        int ln = 0;
        for (int k = 0; k < 3; ++k) x_accumulate_product(DIFF(3 * v + k), DIFF(3 * v + k), k == 0, ws, L, &ln);
        x_copy(ws.at(L.n()), ln, ws.at(kCircSq + 24 * v));
        nsq[v] = ln;
    }
#define SQ(v) ws.at(kCircSq + 24 * (v)), nsq[v]
    // numerator k = |a|^2 (b x c)_k + |b|^2 (c x a)_k + |c|^2 (a x b)_k
    double num[3];
    for (int k = 0; k < 3; ++k) {
        int ln = 0;
        for (int t = 0; t < 3; ++t) {
            if (!x_accumulate_product(SQ(t), CROSS(t, k), t == 0, ws, L, &ln)) {
                *overflow = true;
                return p0;
            }
        }
        num[k] = x_to_double(ws.at(L.n()), ln);
    }
    // den = (a.x (b x c).x + a.y (b x c).y + a.z (b x c).z) * Expansion(2.0)
    int ln = 0;
    for (int k = 0; k < 3; ++k) {
        if (!x_accumulate_product(DIFF(k), CROSS(0, k), k == 0, ws, L, &ln)) {
            *overflow = true;
            return p0;
        }
    }
    if (2 * ln > L.cap_sc || 2 * ln > L.cap_m) {
        *overflow = true;
        return p0;
    }
    Buf r2;
    ws.at(L.t())[0] = 2.0;
    ln = x_mul(ws.at(L.n()), ln, ws.at(L.t()), 1, ws.at(L.sc()), ws.at(L.m0()), ws.at(L.m1()), &r2);
#undef SQ
#undef CROSS
#undef DIFF
    if (ln == 0) return p0;  // coplanar: no circumcentre
    const double d = x_to_double(r2, ln);
    return P3{p0.x + num[0] / d, p0.y + num[1] / d, p0.z + num[2] / d};
}

}  // namespace d3
