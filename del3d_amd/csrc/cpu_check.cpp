// cpu_check.cpp - host build of the per-query tier logic in pred_core.h.
//
// Not part of the product library: built into del3d_amd/libpredcore_host.so
// so the tests can run every tier of the GPU's arithmetic (filter, certified
// small grid, 1-4 limb and any-width exact integers, perturbation walk) on
// the CPU and compare it with the oracle, with a switch that forces a chosen
// tier on every query, which the GPU kernels only reach on rare inputs.
#include <cstdint>
#include <vector>

#include "circ_core.h"
#include "pred2_core.h"
#include "pred_core.h"

using namespace d3;

namespace {

// mode 0: the kernels' tier order; 1: skip T0/T1 (force exact integers,
// 1-4 limbs in registers where possible); 2: force the any-width tier.
int eval_one(uint8_t o, const P3 p_in[5], const int32_t id[5], int mode, int* tier_out) {
    P3 p[5];
    const int np = o == 0 ? 4 : 5;
    for (int k = 0; k < 5; ++k) p[k] = p_in[k < np ? k : 3];
    std::vector<uint32_t> ws(64 * 67);
    const Grid g = grid_of(p, np);
    const int L = limbs_for(g);
    auto exact_orient = [&](const P3* t) -> int {
        if (mode != 2 && L == 1) return orient_public_int<1>(t, g.q);
        if (mode != 2 && L == 2) return orient_public_int<2>(t, g.q);
        if (mode != 2 && L == 3) return orient_public_int<3>(t, g.q);
        if (mode != 2 && L == 4) return orient_public_int<4>(t, g.q);
        return -rt_orient_raw(t[0], t[1], t[2], t[3], g.q, L, ws.data());
    };
    if (o == 0) {
        if (mode == 0) {
            const Filt f = orient_filter(p[0], p[1], p[2], p[3]);
            if (filter_decides(f, kOrientBound)) { *tier_out = 0; return -sgn(f.det); }
            if (t1_certify(p, 4, f.maxd, 15, -358)) { *tier_out = 1; return -sgn(f.det); }
        }
        *tier_out = (mode != 2 && L <= 4) ? 2 : 3;
        return exact_orient(p);
    }
    int raw;
    bool cert = false;
    const Filt f = insphere_filter(p[0], p[1], p[2], p[3], p[4]);
    if (mode == 0 && filter_decides(f, kInSphereBound)) {
        *tier_out = 0;
        return -sgn(f.det);
    }
    if (mode == 0 && t1_certify(p, 5, f.maxd, 8, -214)) {
        *tier_out = 1;
        cert = true;
        raw = -sgn(f.det);
    } else if (mode != 2 && L <= 4) {
        *tier_out = 2;
        raw = -insphere_raw_regs(p, g.q, L);
    } else {
        *tier_out = 3;
        raw = -rt_insphere_raw(p[0], p[1], p[2], p[3], p[4], g.q, L, ws.data());
    }
    if (raw != 0 || o == 2) return raw;
    return perturbation_walk(p, id, [&](int k) {
        P3 t[4];
        subst_tet(p, k, t);
        const Filt ft = orient_filter(t[0], t[1], t[2], t[3]);
        if (cert) return -sgn(ft.det);
        if (mode == 0 && filter_decides(ft, kOrientBound)) return -sgn(ft.det);
        return exact_orient(t);
    });
}

}  // namespace

extern "C" void predcore_host_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                                    const int32_t* idx, int64_t n_q, int8_t* out, int mode,
                                    int8_t* tier) {
    (void)n_pts;
    for (int64_t q = 0; q < n_q; ++q) {
        int32_t id[5];
        P3 p[5];
        for (int k = 0; k < 5; ++k) id[k] = idx[5 * q + k];
        const int np = op[q] == 0 ? 4 : 5;
        for (int k = 0; k < 5; ++k) p[k] = load_p3(xyz, id[k < np ? k : 3]);
        int t = 0;
        out[q] = (int8_t)eval_one(op[q], p, id, mode, &t);
        if (tier) tier[q] = (int8_t)t;
    }
}

// The two T1 certificates of a query's points (np = 4 or 5 points, 3 doubles
// each) against a set summary: out[0] = per-coordinate test, out[1] = set test.
extern "C" void predcore_host_t1(const double* pts, int np, int32_t gmin, int32_t mexp, int* out) {
    P3 p[5];
    for (int k = 0; k < np; ++k) p[k] = P3{pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]};
    const int k = np == 5 ? 8 : 15, qmin = np == 5 ? -214 : -358;
    const double md = max_diff(p, np);
    out[0] = t1_certify(p, np, md, k, qmin) ? 1 : 0;
    out[1] = t1_certify_set(md, k, qmin, SetGrid{gmin, mexp}) ? 1 : 0;
}

// The packed small-grid format of pred_core.h on a set with summary
// (gmin, mexp): returns whether the set qualifies; if it does, words[i] is
// point i packed and back[3i..3i+2] its unpacked coordinates.
extern "C" int predcore_pack_points(const double* xyz, int64_t n, int32_t gmin, int32_t mexp,
                                    uint32_t* words, double* back) {
    const SetGrid g{gmin, mexp};
    if (!set_packable(g)) return 0;
    const double inv = pow2i(-gmin), scale = pow2i(gmin);
    for (int64_t i = 0; i < n; ++i) {
        words[i] = pack_point(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], inv);
        const P3 p = unpack_point(words[i], scale);
        back[3 * i] = p.x;
        back[3 * i + 1] = p.y;
        back[3 * i + 2] = p.z;
    }
    return 1;
}

// The packed kernel's per-query decision on a packable set with summary
// (gmin, mexp): packed_decide on the packed words. Returns 0 (nothing
// written) when the set is not packable; else out[q] = sign.
extern "C" int predcore_packed_batch(const double* xyz, int64_t n_pts, int32_t gmin, int32_t mexp,
                                     const uint8_t* op, const int32_t* idx, int64_t n_q, int8_t* out) {
    if (!set_packable(SetGrid{gmin, mexp})) return 0;
    const double inv = pow2i(-gmin);
    std::vector<uint32_t> words((size_t)n_pts);
    for (int64_t i = 0; i < n_pts; ++i) words[i] = pack_point(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], inv);
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* id = idx + 5 * q;
        const int np = op[q] == DEL3D_OP_ORIENT3D ? 4 : 5;
        uint32_t w[5] = {0, 0, 0, 0, 0};
        for (int k = 0; k < np; ++k) w[k] = words[id[k]];
        int tier = -1;
        bool perturbed = false;
        out[q] = (int8_t)packed_decide(w, op[q], id, tier, perturbed);
    }
    return 1;
}

// Host build of the coplanar predicates (pred2_core.h). mode 0: the full
// evaluation (filters, exact tiers, perturbation); 1: the fast kernel's
// result (9 where it leaves the query to the exact kernel);
// 2: the exact tiers forced (the filters skipped) for orient2d and the raw
// in-circle determinant.
extern "C" void predcore2_host_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                                     const int32_t* idx, int64_t n_q, int8_t* out, int mode) {
    (void)n_pts;
    std::vector<uint32_t> ws((size_t)in_circle_workspace_limbs(kMaxLimbs2));
    for (int64_t q = 0; q < n_q; ++q) {
        const uint8_t o = op[q];
        const int np = points_of_op2(o);
        P3 p[4];
        int32_t id[4];
        for (int k = 0; k < 4; ++k) {
            id[k] = idx[5 * q + k];
            p[k] = k < np ? load_p3(xyz, id[k]) : P3{0.0, 0.0, 0.0};
        }
        int s;
        if (mode == 1) {
            s = pred2_fast(o, p, id);
        } else if (mode == 2 && o == kOp2Orient2dXY) {
            s = orient2d_exact(p[0].x, p[0].y, p[1].x, p[1].y, p[2].x, p[2].y, ws.data());
        } else if (mode == 2 && o == kOp2InCircleUnperturbed) {
            s = in_circle_exact(p[0], p[1], p[2], p[3], ws.data());
        } else {
            s = pred2_eval(o, p, id, ws.data());
        }
        out[q] = (int8_t)s;
    }
}

// Host build of the circumcentres (circ_core.h): lazy or exact, per cell.
static int64_t g_circ_overflows = 0;

// Cells of the last batches that outgrew the small workspace.
extern "C" int64_t circcore_overflows(void) { return g_circ_overflows; }

// The small workspace first, the full one for cells that overflow it (as the
// GPU's two exact passes do).
extern "C" void circcore_host_batch(const double* xyz, const int32_t* cells, int64_t n, int lazy,
                                    double* out) {
    std::vector<double> ws((size_t)kCircFull.total());
    const Buf b{ws.data(), 1};
    for (int64_t c = 0; c < n; ++c) {
        P3 p[4];
        for (int k = 0; k < 4; ++k) p[k] = load_p3(xyz, cells[4 * c + k]);
        double o[3];
        int need = 7;
        if (lazy) need = circ_lazy_pass(p[0], p[1], p[2], p[3], o);
        if (need) {
            bool over = false;
            P3 e = circ_exact(p[0], p[1], p[2], p[3], b, kCircSmall, &over);
            if (over) {
                ++g_circ_overflows;
                e = circ_exact(p[0], p[1], p[2], p[3], b, kCircFull, &over);
            }
            if (need & 1) o[0] = e.x;
            if (need & 2) o[1] = e.y;
            if (need & 4) o[2] = e.z;
        }
        for (int k = 0; k < 3; ++k) out[3 * c + k] = o[k];
    }
}
