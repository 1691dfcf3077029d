// pred_core.h - per-query arithmetic of the batched predicates.
//
// Every function here is __host__ __device__: the GPU kernels call them per
// lane, and a host build of the same header (csrc/cpu_check.cpp) lets the
// tier logic be checked against the oracle without a GPU.
//
// A query is decided by the first tier that can:
//   T0  the reference's semi-static FP64 filter, evaluated in the same
//       operation order with no contraction (src/predicates.cpp:125-144,
//       :149-201), so its accept decisions and signs are the reference's;
//   T1  a certificate that the filter's own FP64 determinant was computed
//       without rounding: all coordinates lie on one power-of-two grid and
//       the differences are small enough that every product and sum of the
//       expression is an integer below 2^53 on that grid. Then sign(det) is
//       the exact sign. Lattice and quantised inputs resolve here, in
//       registers, including the symbolic perturbation;
//   T2  exact integer arithmetic on the query's own power-of-two grid with
//       L = 1..3 32-bit limbs per difference (registers, fully unrolled);
//   T3  the same exact integer arithmetic with any limb count (up to 66,
//       the full double exponent range), on a caller-provided buffer (LDS).
// T2/T3 follow the method of the reference's test oracle (tests/exact.h:
// 161-266): coordinates become integers on their common smallest scale,
// which scales each determinant by a positive power of two. The sign of an
// exact determinant is unique, so every exact tier returns the sign the
// reference's expansion stage (src/predicates.cpp:46-113) returns.
//
// Whatever tier decides, the symbolic perturbation follows the reference's
// walk (src/predicates.cpp:206-232): stable lexicographic order of the five
// argument slots, then slots 4 and 3 tested by index identity against e, d,
// c, b, a in that order, each substitution decided by the public orient3d.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "../../include/del3d_amd.h"

namespace d3 {

#define D3_HD __host__ __device__ inline

// ---- constants (src/predicates.cpp:22-27) ----------------------------------
constexpr double kEps = 1.1102230246251565e-16;  // 2^-53
constexpr double kOrientBound = (7.0 + 56.0 * kEps) * kEps;
constexpr double kInSphereBound = (16.0 + 224.0 * kEps) * kEps;

// Decision tiers, for statistics.
enum Tier : int { kTierFilter = 0, kTierCert = 1, kTierInt = 2, kTierWide = 3 };

struct P3 {
    double x, y, z;
};

D3_HD P3 load_p3(const double* xyz, int32_t i) {
    const double* p = xyz + 3 * (int64_t)i;
    return P3{p[0], p[1], p[2]};
}

D3_HD int sgn(double v) { return (v > 0.0) - (v < 0.0); }

struct Filt {
    double det;   // |a-d, b-d, c-d| (orient3d) or the lifted determinant vs e
    double perm;  // the same expression on absolute values
    double maxd;  // largest |coordinate difference|, used by the T1 certificate
};

// ---- T0: the filters, in the reference's operation order --------------------

// src/predicates.cpp:126-142
D3_HD Filt orient_filter(const P3& a, const P3& b, const P3& c, const P3& d) {
    const double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    const double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    Filt f;
    f.det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
    f.perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) + (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
             (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
    double m = fmax(fmax(fabs(adx), fabs(bdx)), fabs(cdx));
    m = fmax(m, fmax(fmax(fabs(ady), fabs(bdy)), fabs(cdy)));
    f.maxd = fmax(m, fmax(fmax(fabs(adz), fabs(bdz)), fabs(cdz)));
    return f;
}

// The determinants alone, with exactly the filters' operations (so the same
// bits): for queries whose T1 certificate is tried before the filter.
D3_HD double orient_det(const P3& a, const P3& b, const P3& c, const P3& d) {
    const double adx = a.x - d.x, bdx = b.x - d.x, cdx = c.x - d.x;
    const double ady = a.y - d.y, bdy = b.y - d.y, cdy = c.y - d.y;
    const double adz = a.z - d.z, bdz = b.z - d.z, cdz = c.z - d.z;
    const double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
    const double cdxady = cdx * ady, adxcdy = adx * cdy;
    const double adxbdy = adx * bdy, bdxady = bdx * ady;
    return adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
}

D3_HD double insphere_det(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e) {
    const double aex = a.x - e.x, aey = a.y - e.y, aez = a.z - e.z;
    const double bex = b.x - e.x, bey = b.y - e.y, bez = b.z - e.z;
    const double cex = c.x - e.x, cey = c.y - e.y, cez = c.z - e.z;
    const double dex = d.x - e.x, dey = d.y - e.y, dez = d.z - e.z;
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
    return (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
}

// src/predicates.cpp:153-198
D3_HD Filt insphere_filter(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e) {
    const double aex = a.x - e.x, aey = a.y - e.y, aez = a.z - e.z;
    const double bex = b.x - e.x, bey = b.y - e.y, bez = b.z - e.z;
    const double cex = c.x - e.x, cey = c.y - e.y, cez = c.z - e.z;
    const double dex = d.x - e.x, dey = d.y - e.y, dez = d.z - e.z;

    const double aexbey = aex * bey, bexaey = bex * aey;
    const double bexcey = bex * cey, cexbey = cex * bey;
    const double cexdey = cex * dey, dexcey = dex * cey;
    const double dexaey = dex * aey, aexdey = aex * dey;
    const double aexcey = aex * cey, cexaey = cex * aey;
    const double bexdey = bex * dey, dexbey = dex * bey;

    const double ab = aexbey - bexaey;
    const double bc = bexcey - cexbey;
    const double cd = cexdey - dexcey;
    const double da = dexaey - aexdey;
    const double ac = aexcey - cexaey;
    const double bd = bexdey - dexbey;

    const double abc = aez * bc - bez * ac + cez * ab;
    const double bcd = bez * cd - cez * bd + dez * bc;
    const double cda = cez * da + dez * ac + aez * cd;
    const double dab = dez * ab + aez * bd + bez * da;

    const double alift = aex * aex + aey * aey + aez * aez;
    const double blift = bex * bex + bey * bey + bez * bez;
    const double clift = cex * cex + cey * cey + cez * cez;
    const double dlift = dex * dex + dey * dey + dez * dez;

    Filt f;
    f.det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);

    const double az = fabs(aez), bz = fabs(bez), cz = fabs(cez), dz = fabs(dez);
    const double p_ab = fabs(aexbey), p_ba = fabs(bexaey);
    const double p_bc = fabs(bexcey), p_cb = fabs(cexbey);
    const double p_cd = fabs(cexdey), p_dc = fabs(dexcey);
    const double p_da = fabs(dexaey), p_ad = fabs(aexdey);
    const double p_ac = fabs(aexcey), p_ca = fabs(cexaey);
    const double p_bd = fabs(bexdey), p_db = fabs(dexbey);
    f.perm = ((p_cd + p_dc) * bz + (p_db + p_bd) * cz + (p_bc + p_cb) * dz) * alift +
             ((p_da + p_ad) * cz + (p_ac + p_ca) * dz + (p_cd + p_dc) * az) * blift +
             ((p_ab + p_ba) * dz + (p_bd + p_db) * az + (p_da + p_ad) * bz) * clift +
             ((p_bc + p_cb) * az + (p_ca + p_ac) * bz + (p_ab + p_ba) * cz) * dlift;

    double m = fmax(fmax(fabs(aex), fabs(bex)), fmax(fabs(cex), fabs(dex)));
    m = fmax(m, fmax(fmax(fabs(aey), fabs(bey)), fmax(fabs(cey), fabs(dey))));
    f.maxd = fmax(m, fmax(fmax(az, bz), fmax(cz, dz)));
    return f;
}

// The filters' maxd on its own (same differences, same maximum), for callers
// that evaluate the filter first and need maxd only when it is undecided.
D3_HD double max_diff(const P3* p, int np) {
    const P3& r = p[np - 1];
    double m = 0.0;
    for (int i = 0; i < np - 1; ++i) {
        m = fmax(m, fabs(p[i].x - r.x));
        m = fmax(m, fabs(p[i].y - r.y));
        m = fmax(m, fabs(p[i].z - r.z));
    }
    return m;
}

// The reference's accept test: |det| > bound * permanent (src/predicates.cpp:143-144).
D3_HD bool filter_decides(const Filt& f, double bound_const) {
    const double eb = bound_const * f.perm;
    return f.det > eb || -f.det > eb;
}

// ---- T1: certified FP64 evaluation on a small power-of-two grid --------------
//
// Let M be the largest |difference| and q = ilogb(M) - k. If every coordinate
// c is an integer multiple of 2^q with |c| < 2^(52+q), then every difference
// is exact and below 2^(k+1) in units of 2^q. With k = 8 (degree 5, at most
// 72 * 512^5 < 2^53) or k = 15 (degree 3, at most 6 * 65536^3 < 2^53), every
// product and partial sum of the filter's expression is an integer below 2^53
// in units of 2^(deg*q), hence exactly representable, so the FP64 determinant
// is exact. q is bounded so that 2^(deg*q) stays at or above 2^-1074
// (representable) and the values stay far below overflow.
// The tests below combine conditions with & and | rather than && and ||:
// every term is cheap, and branch-free code keeps a wavefront's lanes together.
D3_HD bool on_grid(double c, double inv_unit) {
    const double t = c * inv_unit;  // exact scaling by a power of two unless it underflows
    // an underflowed or fractional t fails the integer test; t == 0 with c != 0
    // means c underflowed entirely, i.e. it is not on the grid
    return (t == trunc(t)) & (fabs(t) < 4503599627370496.0) & ((t != 0.0) | (c == 0.0));
}

// ilogb for the magnitudes T1 sees: the unbiased exponent of a normal
// positive double; zero and subnormals give -2000 (below any accepted q).
D3_HD int exponent_of(double v) {
    const int f = (int)((__builtin_bit_cast(uint64_t, v) >> 52) & 0x7FF);
    return f == 0 ? -2000 : f - 1023;
}

// 2^e for -1022 <= e <= 1023, exactly.
D3_HD double pow2i(int e) { return __builtin_bit_cast(double, (uint64_t)(e + 1023) << 52); }

// The same certificate from a summary of the whole point set: if every
// nonzero coordinate of the set is a multiple of 2^gmin and below 2^(mexp+1)
// in magnitude, then for q <= gmin every coordinate is a multiple of 2^q, and
// mexp <= q + 51 gives |c| < 2^(52+q). One comparison chain per query instead
// of a test per coordinate; when it fails the per-coordinate test still may
// pass (gmin is the set's coarsest common grid, a query's may be coarser).
struct SetGrid {
    int32_t gmin;  // min over nonzero coordinates of the lowest set bit's exponent
    int32_t mexp;  // max over nonzero coordinates of floor(log2 |c|)
};

D3_HD bool t1_certify_set(double maxd, int k, int qmin_allowed, SetGrid g) {
    const int e = exponent_of(maxd);
    const int q = e - k;
    return (maxd == 0.0) | ((q >= qmin_allowed) & (e <= 180) & (q <= g.gmin) & (g.mexp <= q + 51));
}

// deg 5 (insphere): k = 8, q >= -214; deg 3 (orient3d): k = 15, q >= -358
D3_HD bool t1_certify(const P3* p, int np, double maxd, int k, int qmin_allowed) {
    const int e = exponent_of(maxd);
    const int q = e - k;
    const bool in_range = (q >= qmin_allowed) & (e <= 180);
    const double inv = pow2i(in_range ? -q : 0);  // -q in [k-180, -qmin]: representable
    bool ok = in_range;
    for (int i = 0; i < np; ++i)
        ok = ok & on_grid(p[i].x, inv) & on_grid(p[i].y, inv) & on_grid(p[i].z, inv);
    // maxd == 0: all points identical, every difference is exactly 0
    return (maxd == 0.0) | ok;
}

// Whether a set's summary allows the set-level certificate at all: it needs
// some q with mexp - 51 <= q <= gmin. When it does, the fast kernels try the
// certificate before the filter (a certified determinant's sign is the exact
// sign, so the filter's error bound need not be formed); sets of arbitrary
// doubles fail this once per task and keep the filter-first order.
D3_HD bool set_certifiable(SetGrid g) { return g.mexp - g.gmin <= 51; }

// Whether every coordinate of a set is exactly a normal float (or zero): its
// nonzero values span at most 24 significant bits (mexp - gmin <= 23), none
// below 2^-126 (so no float subnormal, whatever the device's float denormal
// mode) and none at 2^128 or above. An all-zero set passes too.
D3_HD bool set_f32_exact(SetGrid g) { return g.gmin >= -126 && g.mexp <= 127 && g.mexp - g.gmin <= 23; }

// ---- the perturbation walk (src/predicates.cpp:211-231) ---------------------

// (x, y, z) lexicographic <, with IEEE comparisons (so -0.0 == +0.0), as
// src/predicates.cpp:117-121 (for finite coordinates).
D3_HD bool lex_less(const P3& p, const P3& q) {
    return (p.x < q.x) | ((p.x == q.x) & ((p.y < q.y) | ((p.y == q.y) & (p.z < q.z))));
}

// Slots 0..4 are (a, b, c, d, e). The walk needs the slots at sorted ranks 4
// and 3 of the stable lexicographic order; with ties broken by slot, the rank
// of slot s is the number of slots that come before it.
// less(t, s): slot t's point is lexicographically below slot s's.
template <class Less>
D3_HD void top_two_slots_by(Less&& less, int& s4, int& s3) {
    s4 = 0;
    s3 = 0;
#pragma unroll
    for (int s = 0; s < 5; ++s) {
        int r = 0;
#pragma unroll
        for (int t = 0; t < 5; ++t) {
            if (t == s) continue;
            const bool before = less(t, s) || (t < s && !less(s, t));
            r += before ? 1 : 0;
        }
        if (r == 4) s4 = s;
        if (r == 3) s3 = s;
    }
}

D3_HD void top_two_slots(const P3 p[5], int& s4, int& s3) {
    top_two_slots_by([&](int t, int s) { return lex_less(p[t], p[s]); }, s4, s3);
}

// `orient_subst(k)` returns the public orient3d sign of the tetrahedron
// (a, b, c, d) with slot k (0..3) replaced by e; it is only called for the
// substitutions the walk reaches.
template <class F>
D3_HD int perturbation_walk_from(int s4, int s3, const int32_t id[5], F&& orient_subst) {
#pragma unroll
    for (int it = 0; it < 2; ++it) {
        // a select chain rather than id[slot]: a runtime index would put the
        // array in private (scratch) memory on the GPU
        const int sl = it == 0 ? s4 : s3;
        const int32_t v = sl == 0 ? id[0] : sl == 1 ? id[1] : sl == 2 ? id[2] : sl == 3 ? id[3] : id[4];
        if (v == id[4]) return -1;
        // slots d, c, b, a in that order; with distinct indices exactly one
        // matches, so the common case evaluates one substitution, in one
        // place for every lane; aliased slots continue in order below
        const int k = v == id[3] ? 3 : v == id[2] ? 2 : v == id[1] ? 1 : v == id[0] ? 0 : -1;
        if (k >= 0) {
            int o = orient_subst(k);
            if (o != 0) return o;
            for (int k2 = k - 1; k2 >= 0; --k2) {
                const int32_t w = k2 == 2 ? id[2] : k2 == 1 ? id[1] : id[0];
                if (v == w && (o = orient_subst(k2)) != 0) return o;
            }
        }
    }
    return -1;
}

template <class F>
D3_HD int perturbation_walk(const P3 p[5], const int32_t id[5], F&& orient_subst) {
    int s4, s3;
    top_two_slots(p, s4, s3);
    return perturbation_walk_from(s4, s3, id, orient_subst);
}

D3_HD void subst_tet(const P3 p[5], int k, P3 t[4]) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const bool sub = i == k;
        t[i] = P3{sub ? p[4].x : p[i].x, sub ? p[4].y : p[i].y, sub ? p[4].z : p[i].z};
    }
}

// ---- packed small-grid sets ---------------------------------------------------
// A set whose nonzero coordinates are all c = k * 2^gmin with |k| <= 255 (its
// values span at most 8 significant bits: mexp - gmin <= 7), with
// -206 <= gmin <= 171, is stored as one 32-bit word per point: three 10-bit
// fields k + 512, x in bits 20-29, y in 10-19, z in 0-9, so that comparing
// words is comparing points lexicographically. -0.0 is stored as 0.0: no
// fast tier depends on the sign of a zero coordinate (a zero operand adds or
// multiplies to the same value either way, only the sign of a zero result can
// differ, and every test on such a result, the filter's comparisons, the
// grid test, sgn() and the lexicographic order, treats -0.0 as 0.0).
// An all-zero set does not qualify.
D3_HD bool set_packable(SetGrid g) {
    return g.gmin <= g.mexp && g.gmin >= -206 && g.gmin <= 171 && g.mexp - g.gmin <= 7;
}

// inv_scale = 2^-gmin
D3_HD uint32_t pack_point(double x, double y, double z, double inv_scale) {
    const uint32_t fx = (uint32_t)((int32_t)(x * inv_scale) + 512);
    const uint32_t fy = (uint32_t)((int32_t)(y * inv_scale) + 512);
    const uint32_t fz = (uint32_t)((int32_t)(z * inv_scale) + 512);
    return (fx << 20) | (fy << 10) | fz;
}
// scale = 2^gmin: the set's doubles (zeros as +0.0)
D3_HD P3 unpack_point(uint32_t w, double scale) {
    return P3{(double)((int32_t)((w >> 20) & 1023u) - 512) * scale,
              (double)((int32_t)((w >> 10) & 1023u) - 512) * scale,
              (double)((int32_t)(w & 1023u) - 512) * scale};
}

// The difference a - b of two packed points in grid units (exact doubles).
struct DV {
    double x, y, z;
};
D3_HD double pfield_diff(uint32_t a, uint32_t b, int shift) {
    return (double)((int32_t)((a >> shift) & 1023u) - (int32_t)((b >> shift) & 1023u));
}
D3_HD DV pdiff(uint32_t a, uint32_t b) { return DV{pfield_diff(a, b, 20), pfield_diff(a, b, 10), pfield_diff(a, b, 0)}; }
D3_HD DV dv_sub(const DV& a, const DV& b) { return DV{a.x - b.x, a.y - b.y, a.z - b.z}; }

// orient_det and insphere_det on the differences to the last point.
D3_HD double orient_det_dv(const DV& ad, const DV& bd, const DV& cd) {
    const double bdxcdy = bd.x * cd.y, cdxbdy = cd.x * bd.y;
    const double cdxady = cd.x * ad.y, adxcdy = ad.x * cd.y;
    const double adxbdy = ad.x * bd.y, bdxady = bd.x * ad.y;
    return ad.z * (bdxcdy - cdxbdy) + bd.z * (cdxady - adxcdy) + cd.z * (adxbdy - bdxady);
}
D3_HD double insphere_det_dv(const DV& ae, const DV& be, const DV& ce, const DV& de) {
    const double ab = ae.x * be.y - be.x * ae.y;
    const double bc = be.x * ce.y - ce.x * be.y;
    const double cd = ce.x * de.y - de.x * ce.y;
    const double da = de.x * ae.y - ae.x * de.y;
    const double ac = ae.x * ce.y - ce.x * ae.y;
    const double bd = be.x * de.y - de.x * be.y;
    const double abc = ae.z * bc - be.z * ac + ce.z * ab;
    const double bcd = be.z * cd - ce.z * bd + de.z * bc;
    const double cda = ce.z * da + de.z * ac + ae.z * cd;
    const double dab = de.z * ab + ae.z * bd + be.z * da;
    const double alift = ae.x * ae.x + ae.y * ae.y + ae.z * ae.z;
    const double blift = be.x * be.x + be.y * be.y + be.z * be.z;
    const double clift = ce.x * ce.x + ce.y * ce.y + ce.z * ce.z;
    const double dlift = de.x * de.x + de.y * de.y + de.z * de.z;
    return (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
}

// The fast tiers of one query of a packed set, on the integer grid: the
// certified branch of the set-level certificate, which is what the generic
// path takes for every query of these sets. With |k| <= 255 every difference
// is below 512 grid units, so with -206 <= gmin <= 171 the orient3d (k = 15)
// and in_sphere (k = 8) certificates hold for every query; a certified
// determinant is exact, and on the grid it is the set's determinant divided
// by a positive power of two, so its sign is the same, as is every
// substituted orient3d of the perturbation and the points' lexicographic
// order. w: the query's packed points (w[4] for in_sphere) of a packable set.
D3_HD int packed_decide(const uint32_t w[5], uint8_t o, const int32_t id[5], int& tier, bool& perturbed) {
    tier = kTierCert;
    if (o == DEL3D_OP_ORIENT3D)
        return -sgn(orient_det_dv(pdiff(w[0], w[3]), pdiff(w[1], w[3]), pdiff(w[2], w[3])));
    const DV d0 = pdiff(w[0], w[4]), d1 = pdiff(w[1], w[4]), d2 = pdiff(w[2], w[4]), d3 = pdiff(w[3], w[4]);
    int sign = -sgn(insphere_det_dv(d0, d1, d2, d3));
    if (sign == 0 && o == DEL3D_OP_IN_SPHERE) {
        perturbed = true;
        // the slots of ranks 4 and 3 in the stable order (ties by slot): the
        // two largest of (w[s], s), kept while scanning s upwards
        int s4 = 0, s3 = 0;
        uint32_t w4 = w[0], w3 = 0;
#pragma unroll
        for (int s = 1; s < 5; ++s) {
            const bool top = w[s] >= w4;
            const bool second = !top && (s == 1 || w[s] >= w3);
            s3 = top ? s4 : second ? s : s3;
            w3 = top ? w4 : second ? w[s] : w3;
            s4 = top ? s : s4;
            w4 = top ? w[s] : w4;
        }
        sign = perturbation_walk_from(s4, s3, id, [&](int k) {
            // slot k replaced by e; differences to the fourth point, which is
            // e for k = 3 and d otherwise: x - d = (x - e) - (d - e)
            const bool kd = k != 3;
            const DV base{kd ? d3.x : 0.0, kd ? d3.y : 0.0, kd ? d3.z : 0.0};
            const DV nd{-d3.x, -d3.y, -d3.z};
            const DV a = k == 0 ? nd : dv_sub(d0, base);
            const DV b = k == 1 ? nd : dv_sub(d1, base);
            const DV c = k == 2 ? nd : dv_sub(d2, base);
            return -sgn(orient_det_dv(a, b, c));
        });
    }
    return sign;
}

// ---- exact integer arithmetic: two's complement, 32-bit limbs --------------

template <int N>
struct Limbs {
    uint32_t v[N];
};

template <int N>
D3_HD Limbs<N> li_sub(const Limbs<N>& a, const Limbs<N>& b) {
    Limbs<N> r;
    uint32_t borrow = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint64_t t = (uint64_t)a.v[i] - (uint64_t)b.v[i] - borrow;
        r.v[i] = (uint32_t)t;
        borrow = (uint32_t)(t >> 63);
    }
    return r;
}

template <int N>
D3_HD Limbs<N> li_add(const Limbs<N>& a, const Limbs<N>& b) {
    Limbs<N> r;
    uint32_t carry = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint64_t t = (uint64_t)a.v[i] + (uint64_t)b.v[i] + carry;
        r.v[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
    return r;
}

// r = neg ? -a : a
template <int N>
D3_HD Limbs<N> li_cneg(const Limbs<N>& a, uint32_t neg) {
    Limbs<N> r;
    const uint32_t mask = 0u - neg;
    uint32_t carry = neg;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        const uint64_t t = (uint64_t)(a.v[i] ^ mask) + carry;
        r.v[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
    return r;
}

// Signed product; the result width N+M holds any product of an N- and an
// M-limb value.
template <int N, int M>
D3_HD Limbs<N + M> li_mul(const Limbs<N>& a, const Limbs<M>& b) {
    const uint32_t sa = a.v[N - 1] >> 31, sb = b.v[M - 1] >> 31;
    const Limbs<N> ua = li_cneg(a, sa);
    const Limbs<M> ub = li_cneg(b, sb);
    Limbs<N + M> r;
#pragma unroll
    for (int i = 0; i < N + M; ++i) r.v[i] = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) {
        uint64_t carry = 0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const uint64_t t = (uint64_t)ua.v[i] * ub.v[j] + r.v[i + j] + carry;
            r.v[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r.v[i + M] = (uint32_t)carry;
    }
    return li_cneg(r, sa ^ sb);
}

template <int N>
D3_HD int li_sign(const Limbs<N>& a) {
    if (a.v[N - 1] >> 31) return -1;
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < N; ++i) any |= a.v[i];
    return any ? 1 : 0;
}

// ---- the query's power-of-two grid -----------------------------------------

struct Grid {
    int q;    // every coordinate is an integer multiple of 2^q
    int top;  // every |coordinate| < 2^(top+1)
    bool any; // some coordinate is nonzero
};

struct Bits {
    uint64_t m;  // significand, with the implicit bit
    int e;       // value = m * 2^e
    bool neg;
};

D3_HD Bits decompose(double c) {
    const uint64_t b = __builtin_bit_cast(uint64_t, c);
    const int ef = (int)((b >> 52) & 0x7FF);
    Bits r;
    r.m = b & 0xFFFFFFFFFFFFFull;
    r.neg = (b >> 63) != 0;
    if (ef == 0) {
        r.e = -1074;
    } else {
        r.m |= (1ull << 52);
        r.e = ef - 1075;
    }
    return r;
}

D3_HD void grid_add(Grid& g, double c) {
    const Bits d = decompose(c);
    if (d.m == 0) return;
    const int lsb = d.e + __builtin_ctzll(d.m);
    const int msb = d.e + 63 - __builtin_clzll(d.m);
    if (!g.any) {
        g.q = lsb;
        g.top = msb;
        g.any = true;
    } else {
        g.q = lsb < g.q ? lsb : g.q;
        g.top = msb > g.top ? msb : g.top;
    }
}

D3_HD Grid grid_of(const P3* p, int np) {
    Grid g{0, 0, false};
    for (int i = 0; i < np; ++i) {
        grid_add(g, p[i].x);
        grid_add(g, p[i].y);
        grid_add(g, p[i].z);
    }
    return g;
}

// Limbs per difference: |coordinate| < 2^B with B = top + 1 - q, so
// |difference| < 2^(B+1); keeping 8 spare bits below the sign bit makes every
// product of the determinants fit its natural width (degree k in k*L limbs).
D3_HD int limbs_for(const Grid& g) {
    if (!g.any) return 1;
    const int B = g.top + 1 - g.q;
    return (B + 9 + 31) / 32;
}

// bits [32*i, 32*i+32) of m * 2^s (s may be negative when m has trailing
// zeros below the grid, which are then the only bits dropped)
D3_HD uint32_t shifted_word(uint64_t m, int s, int i) {
    const int pos = 32 * i - s;  // bit of m that lands on bit 0 of word i
    if (pos >= 64 || pos <= -32) return 0u;
    if (pos >= 0) return (uint32_t)(m >> pos);
    return (uint32_t)(m << (-pos));
}

template <int L>
D3_HD Limbs<L> to_limbs(double c, int q) {
    const Bits d = decompose(c);
    Limbs<L> r;
#pragma unroll
    for (int i = 0; i < L; ++i) r.v[i] = d.m ? shifted_word(d.m, d.e - q, i) : 0u;
    return li_cneg(r, d.neg && d.m ? 1u : 0u);
}

template <int L>
struct LP3 {
    Limbs<L> x, y, z;
};

template <int L>
D3_HD LP3<L> to_lp3(const P3& p, int q) {
    return LP3<L>{to_limbs<L>(p.x, q), to_limbs<L>(p.y, q), to_limbs<L>(p.z, q)};
}

// sign of |a-d, b-d, c-d| (the computed form; the public orient3d is its negation)
template <int L>
D3_HD int int_orient_raw(const LP3<L>& a, const LP3<L>& b, const LP3<L>& c, const LP3<L>& d) {
    const Limbs<L> adx = li_sub(a.x, d.x), ady = li_sub(a.y, d.y), adz = li_sub(a.z, d.z);
    const Limbs<L> bdx = li_sub(b.x, d.x), bdy = li_sub(b.y, d.y), bdz = li_sub(b.z, d.z);
    const Limbs<L> cdx = li_sub(c.x, d.x), cdy = li_sub(c.y, d.y), cdz = li_sub(c.z, d.z);
    const Limbs<2 * L> m1 = li_sub(li_mul(bdx, cdy), li_mul(cdx, bdy));
    const Limbs<2 * L> m2 = li_sub(li_mul(cdx, ady), li_mul(adx, cdy));
    const Limbs<2 * L> m3 = li_sub(li_mul(adx, bdy), li_mul(bdx, ady));
    const Limbs<3 * L> det = li_add(li_add(li_mul(adz, m1), li_mul(bdz, m2)), li_mul(cdz, m3));
    return li_sign(det);
}

// sign of the lifted determinant relative to e (the in_sphere computed form;
// the public in_sphere_unperturbed is its negation)
template <int L>
D3_HD int int_insphere_raw(const LP3<L>& a, const LP3<L>& b, const LP3<L>& c, const LP3<L>& d,
                           const LP3<L>& e) {
    const Limbs<L> ax = li_sub(a.x, e.x), ay = li_sub(a.y, e.y), az = li_sub(a.z, e.z);
    const Limbs<L> bx = li_sub(b.x, e.x), by = li_sub(b.y, e.y), bz = li_sub(b.z, e.z);
    const Limbs<L> cx = li_sub(c.x, e.x), cy = li_sub(c.y, e.y), cz = li_sub(c.z, e.z);
    const Limbs<L> dx = li_sub(d.x, e.x), dy = li_sub(d.y, e.y), dz = li_sub(d.z, e.z);

    const Limbs<2 * L> ab = li_sub(li_mul(ax, by), li_mul(bx, ay));
    const Limbs<2 * L> bc = li_sub(li_mul(bx, cy), li_mul(cx, by));
    const Limbs<2 * L> cd = li_sub(li_mul(cx, dy), li_mul(dx, cy));
    const Limbs<2 * L> da = li_sub(li_mul(dx, ay), li_mul(ax, dy));
    const Limbs<2 * L> ac = li_sub(li_mul(ax, cy), li_mul(cx, ay));
    const Limbs<2 * L> bd = li_sub(li_mul(bx, dy), li_mul(dx, by));

    const Limbs<3 * L> abc = li_add(li_sub(li_mul(az, bc), li_mul(bz, ac)), li_mul(cz, ab));
    const Limbs<3 * L> bcd = li_add(li_sub(li_mul(bz, cd), li_mul(cz, bd)), li_mul(dz, bc));
    const Limbs<3 * L> cda = li_add(li_add(li_mul(cz, da), li_mul(dz, ac)), li_mul(az, cd));
    const Limbs<3 * L> dab = li_add(li_add(li_mul(dz, ab), li_mul(az, bd)), li_mul(bz, da));

    const Limbs<2 * L> la = li_add(li_add(li_mul(ax, ax), li_mul(ay, ay)), li_mul(az, az));
    const Limbs<2 * L> lb = li_add(li_add(li_mul(bx, bx), li_mul(by, by)), li_mul(bz, bz));
    const Limbs<2 * L> lc = li_add(li_add(li_mul(cx, cx), li_mul(cy, cy)), li_mul(cz, cz));
    const Limbs<2 * L> ld = li_add(li_add(li_mul(dx, dx), li_mul(dy, dy)), li_mul(dz, dz));

    const Limbs<5 * L> det = li_add(li_sub(li_mul(ld, abc), li_mul(lc, dab)),
                                    li_sub(li_mul(lb, cda), li_mul(la, bcd)));
    return li_sign(det);
}

// The raw in_sphere sign in registers for 1-4 limbs per difference (the caller
// checks L <= 4).
D3_HD int insphere_raw_regs(const P3 p[5], int q, int L) {
    switch (L) {
    case 1:
        return int_insphere_raw<1>(to_lp3<1>(p[0], q), to_lp3<1>(p[1], q), to_lp3<1>(p[2], q), to_lp3<1>(p[3], q),
                                   to_lp3<1>(p[4], q));
    case 2:
        return int_insphere_raw<2>(to_lp3<2>(p[0], q), to_lp3<2>(p[1], q), to_lp3<2>(p[2], q), to_lp3<2>(p[3], q),
                                   to_lp3<2>(p[4], q));
    case 3:
        return int_insphere_raw<3>(to_lp3<3>(p[0], q), to_lp3<3>(p[1], q), to_lp3<3>(p[2], q), to_lp3<3>(p[3], q),
                                   to_lp3<3>(p[4], q));
    default:
        return int_insphere_raw<4>(to_lp3<4>(p[0], q), to_lp3<4>(p[1], q), to_lp3<4>(p[2], q), to_lp3<4>(p[3], q),
                                   to_lp3<4>(p[4], q));
    }
}

// ---- runtime-width exact arithmetic (T3), on a caller buffer ----------------
// Values are two's complement arrays of n limbs. The buffer is a bump arena.

D3_HD void rt_to_limbs(double c, int q, uint32_t* out, int L) {
    const Bits d = decompose(c);
    uint32_t carry = (d.neg && d.m) ? 1u : 0u;
    const uint32_t mask = 0u - carry;
    for (int i = 0; i < L; ++i) {
        const uint32_t w = d.m ? shifted_word(d.m, d.e - q, i) : 0u;
        const uint64_t t = (uint64_t)(w ^ mask) + carry;
        out[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
}

// r = a - b (sub != 0) or a + b, all n limbs (r may alias a or b)
D3_HD void rt_addsub(uint32_t* r, const uint32_t* a, const uint32_t* b, int n, int sub) {
    const uint32_t mask = sub ? 0xFFFFFFFFu : 0u;
    uint32_t carry = sub ? 1u : 0u;
    for (int i = 0; i < n; ++i) {
        const uint64_t t = (uint64_t)a[i] + (uint64_t)(b[i] ^ mask) + carry;
        r[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
}

D3_HD void rt_cneg(uint32_t* r, const uint32_t* a, int n, uint32_t neg) {
    const uint32_t mask = 0u - neg;
    uint32_t carry = neg;
    for (int i = 0; i < n; ++i) {
        const uint64_t t = (uint64_t)(a[i] ^ mask) + carry;
        r[i] = (uint32_t)t;
        carry = (uint32_t)(t >> 32);
    }
}

// r[na+nb] = a[na] * b[nb], signed; tmp holds na+nb limbs. r must not alias
// a, b or tmp.
D3_HD void rt_mul(uint32_t* r, const uint32_t* a, int na, const uint32_t* b, int nb, uint32_t* tmp) {
    const uint32_t sa = a[na - 1] >> 31, sb = b[nb - 1] >> 31;
    uint32_t* ua = tmp;
    uint32_t* ub = tmp + na;
    rt_cneg(ua, a, na, sa);
    rt_cneg(ub, b, nb, sb);
    for (int i = 0; i < na + nb; ++i) r[i] = 0;
    for (int i = 0; i < na; ++i) {
        uint64_t carry = 0;
        for (int j = 0; j < nb; ++j) {
            const uint64_t t = (uint64_t)ua[i] * ub[j] + r[i + j] + carry;
            r[i + j] = (uint32_t)t;
            carry = t >> 32;
        }
        r[i + nb] = (uint32_t)carry;
    }
    rt_cneg(r, r, na + nb, sa ^ sb);
}

D3_HD int rt_sign(const uint32_t* a, int n) {
    if (a[n - 1] >> 31) return -1;
    uint32_t any = 0;
    for (int i = 0; i < n; ++i) any |= a[i];
    return any ? 1 : 0;
}

// acc (n limbs) += or -= x*y, with x of nx limbs and y of ny limbs, nx+ny == n
D3_HD void rt_mac(uint32_t* acc, const uint32_t* x, int nx, const uint32_t* y, int ny, int sub,
                  uint32_t* prod, uint32_t* tmp) {
    rt_mul(prod, x, nx, y, ny, tmp);
    rt_addsub(acc, acc, prod, nx + ny, sub);
}

// Buffer limbs the wide tier needs for limb count L (insphere; orient3d needs less).
D3_HD int rt_workspace_limbs(int L) { return 64 * L; }

// Wide-tier orient3d: sign of |a-d, b-d, c-d|. ws holds rt_workspace_limbs(L).
D3_HD int rt_orient_raw(const P3& a, const P3& b, const P3& c, const P3& d, int q, int L,
                        uint32_t* ws) {
    uint32_t* D = ws;           // 9 differences, L each
    uint32_t* m = D + 9 * L;    // 3 minors, 2L each
    uint32_t* det = m + 6 * L;  // 3L
    uint32_t* prod = det + 3 * L;
    uint32_t* tmp = prod + 3 * L;
    uint32_t* t2 = tmp + 3 * L;
    const P3* pt[3] = {&a, &b, &c};
    for (int i = 0; i < 3; ++i) {
        rt_to_limbs(pt[i]->x, q, D + (3 * i + 0) * L, L);
        rt_to_limbs(pt[i]->y, q, D + (3 * i + 1) * L, L);
        rt_to_limbs(pt[i]->z, q, D + (3 * i + 2) * L, L);
    }
    rt_to_limbs(d.x, q, t2, L);
    rt_to_limbs(d.y, q, t2 + L, L);
    rt_to_limbs(d.z, q, t2 + 2 * L, L);
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) rt_addsub(D + (3 * i + k) * L, D + (3 * i + k) * L, t2 + k * L, L, 1);
    // X(i) = D+3iL (x), Y(i) = D+(3i+1)L, Z(i) = D+(3i+2)L with i = a,b,c
#define RX(i) (D + (3 * (i) + 0) * L)
#define RY(i) (D + (3 * (i) + 1) * L)
#define RZ(i) (D + (3 * (i) + 2) * L)
    // m1 = bdx*cdy - cdx*bdy ; m2 = cdx*ady - adx*cdy ; m3 = adx*bdy - bdx*ady
    const int mi[3][4] = {{1, 2, 2, 1}, {2, 0, 0, 2}, {0, 1, 1, 0}};
    for (int k = 0; k < 3; ++k) {
        uint32_t* mk = m + 2 * k * L;
        rt_mul(mk, RX(mi[k][0]), L, RY(mi[k][1]), L, tmp);
        rt_mac(mk, RX(mi[k][2]), L, RY(mi[k][3]), L, 1, prod, tmp);
    }
    for (int i = 0; i < 3 * L; ++i) det[i] = 0;
    for (int k = 0; k < 3; ++k) rt_mac(det, RZ(k), L, m + 2 * k * L, 2 * L, 0, prod, tmp);
#undef RX
#undef RY
#undef RZ
    return rt_sign(det, 3 * L);
}

// Wide-tier lifted determinant relative to e (in_sphere computed form).
D3_HD int rt_insphere_raw(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e, int q,
                          int L, uint32_t* ws) {
    uint32_t* D = ws;              // 12 differences (a,b,c,d) x (x,y,z), L each
    uint32_t* mn = D + 12 * L;     // 6 minors ab bc cd da ac bd, 2L each
    uint32_t* cf = mn + 12 * L;    // 4 cofactors abc bcd cda dab, 3L each
    uint32_t* lf = cf + 12 * L;    // 4 lifts, 2L each
    uint32_t* det = lf + 8 * L;    // 5L
    uint32_t* prod = det + 5 * L;  // 5L
    uint32_t* tmp = prod + 5 * L;  // 5L
    uint32_t* ev = tmp + 5 * L;    // 3L: e's coordinates (total 64L)
    const P3* pt[4] = {&a, &b, &c, &d};
    rt_to_limbs(e.x, q, ev, L);
    rt_to_limbs(e.y, q, ev + L, L);
    rt_to_limbs(e.z, q, ev + 2 * L, L);
    for (int i = 0; i < 4; ++i) {
        rt_to_limbs(pt[i]->x, q, D + (3 * i + 0) * L, L);
        rt_to_limbs(pt[i]->y, q, D + (3 * i + 1) * L, L);
        rt_to_limbs(pt[i]->z, q, D + (3 * i + 2) * L, L);
        for (int k = 0; k < 3; ++k) rt_addsub(D + (3 * i + k) * L, D + (3 * i + k) * L, ev + k * L, L, 1);
    }
#define RX(i) (D + (3 * (i) + 0) * L)
#define RY(i) (D + (3 * (i) + 1) * L)
#define RZ(i) (D + (3 * (i) + 2) * L)
#define MN(k) (mn + 2 * (k) * L)
    // minors: ab(0,1) bc(1,2) cd(2,3) da(3,0) ac(0,2) bd(1,3): p.x*q.y - q.x*p.y
    const int mp[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 2}, {1, 3}};
    for (int k = 0; k < 6; ++k) {
        rt_mul(MN(k), RX(mp[k][0]), L, RY(mp[k][1]), L, tmp);
        rt_mac(MN(k), RX(mp[k][1]), L, RY(mp[k][0]), L, 1, prod, tmp);
    }
    // cofactors (src/predicates.cpp:165-168):
    //   abc = az*bc - bz*ac + cz*ab ; bcd = bz*cd - cz*bd + dz*bc
    //   cda = cz*da + dz*ac + az*cd ; dab = dz*ab + az*bd + bz*da
    const int cz_[4][3] = {{0, 1, 2}, {1, 2, 3}, {2, 3, 0}, {3, 0, 1}};   // z-point per term
    const int cm_[4][3] = {{1, 4, 0}, {2, 5, 1}, {3, 4, 2}, {0, 5, 3}};   // minor per term
    const int cs_[4][3] = {{0, 1, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0}};   // 1 = subtract
    for (int k = 0; k < 4; ++k) {
        uint32_t* ck = cf + 3 * k * L;
        for (int i = 0; i < 3 * L; ++i) ck[i] = 0;
        for (int t = 0; t < 3; ++t) rt_mac(ck, RZ(cz_[k][t]), L, MN(cm_[k][t]), 2 * L, cs_[k][t], prod, tmp);
    }
    for (int k = 0; k < 4; ++k) {
        uint32_t* lk = lf + 2 * k * L;
        for (int i = 0; i < 2 * L; ++i) lk[i] = 0;
        rt_mac(lk, RX(k), L, RX(k), L, 0, prod, tmp);
        rt_mac(lk, RY(k), L, RY(k), L, 0, prod, tmp);
        rt_mac(lk, RZ(k), L, RZ(k), L, 0, prod, tmp);
    }
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd)
    for (int i = 0; i < 5 * L; ++i) det[i] = 0;
    rt_mac(det, lf + 6 * L, 2 * L, cf + 0 * L, 3 * L, 0, prod, tmp);
    rt_mac(det, lf + 4 * L, 2 * L, cf + 9 * L, 3 * L, 1, prod, tmp);
    rt_mac(det, lf + 2 * L, 2 * L, cf + 6 * L, 3 * L, 0, prod, tmp);
    rt_mac(det, lf + 0 * L, 2 * L, cf + 3 * L, 3 * L, 1, prod, tmp);
#undef RX
#undef RY
#undef RZ
#undef MN
    return rt_sign(det, 5 * L);
}

// ---- tiered orient3d for a perturbation substitution or a plain query -------
// Returns the public orient3d sign, using the cheapest tier that is exact.
// `cert_grid` says the five points already passed the degree-5 certificate,
// which implies the degree-3 one for any four of them.
template <int L>
D3_HD int orient_public_int(const P3 t[4], int q) {
    return -int_orient_raw<L>(to_lp3<L>(t[0], q), to_lp3<L>(t[1], q), to_lp3<L>(t[2], q),
                              to_lp3<L>(t[3], q));
}

}  // namespace d3
