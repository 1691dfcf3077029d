// host_scalar.cpp - the scalar entry points (one orient3d / in_sphere /
// coplanar predicate / circumcentre per call, the C++ drop-in's calls),
// evaluated on the host with the per-query arithmetic the kernels run
// (pred_core.h, pred2_core.h, circ_core.h), compiled here for the CPU with
// the same floating-point rules (no contraction, no fast-math). A kernel
// launch per call would cost a 10-20 us device round trip for tens of
// nanoseconds of work; the reference's Delaunay driver makes about 47 such
// calls per inserted point (src/triangulation.cpp:430,476,596,600,793).
// Batches still go to the GPU.
#include <vector>

#include "../../include/del3d_amd.h"
#include "circ_core.h"
#include "host_scalar.h"
#include "pred2_core.h"
#include "pred_core.h"

namespace d3 {
namespace {

constexpr int kWideLimbs = 67;  // > limbs_for() of any finite input (66)

// Workspaces of the any-width exact tiers, one per calling thread (the
// entry points are reentrant, as the reference's are).
uint32_t* wide_workspace() {
    thread_local std::vector<uint32_t> ws(64 * kWideLimbs);
    return ws.data();
}

uint32_t* circle_workspace() {
    thread_local std::vector<uint32_t> ws((size_t)in_circle_workspace_limbs(kMaxLimbs2));
    return ws.data();
}

double* circumcenter_workspace() {
    thread_local std::vector<double> ws((size_t)kCircFull.total());
    return ws.data();
}

}  // namespace

// The tier order of the kernels' per-query path: the reference's filter
// (src/predicates.cpp:125-201), then exact integers on the query's own
// power-of-two grid (1-4 limbs unrolled, else any width), then for in_sphere
// the symbolic perturbation (src/predicates.cpp:206-232), whose substituted
// orient3d calls take the same route.
int host_pred3(uint8_t op, const P3 pin[5], const int32_t id[5]) {
    P3 p[5];
    const int np = op == DEL3D_OP_ORIENT3D ? 4 : 5;
    for (int k = 0; k < 5; ++k) p[k] = pin[k < np ? k : 3];
    const Grid g = grid_of(p, np);
    const int L = limbs_for(g);
    uint32_t* ws = wide_workspace();
    auto orient_pub = [&](const P3* t) -> int {
        const Filt f = orient_filter(t[0], t[1], t[2], t[3]);
        if (filter_decides(f, kOrientBound)) return -sgn(f.det);
        if (L == 1) return orient_public_int<1>(t, g.q);
        if (L == 2) return orient_public_int<2>(t, g.q);
        if (L == 3) return orient_public_int<3>(t, g.q);
        if (L == 4) return orient_public_int<4>(t, g.q);
        return -rt_orient_raw(t[0], t[1], t[2], t[3], g.q, L, ws);
    };
    if (op == DEL3D_OP_ORIENT3D) return orient_pub(p);
    const Filt f = insphere_filter(p[0], p[1], p[2], p[3], p[4]);
    int s;
    if (filter_decides(f, kInSphereBound)) {
        s = -sgn(f.det);
    } else if (L <= 4) {
        s = -insphere_raw_regs(p, g.q, L);
    } else {
        s = -rt_insphere_raw(p[0], p[1], p[2], p[3], p[4], g.q, L, ws);
    }
    if (s != 0 || op != DEL3D_OP_IN_SPHERE) return s;
    return perturbation_walk(p, id, [&](int k) {
        P3 t[4];
        subst_tet(p, k, t);
        return orient_pub(t);
    });
}

int host_pred2(uint8_t op, const P3 p[4], const int32_t id[4]) {
    return pred2_eval(op, p, id, circle_workspace());
}

void host_circumcenter(const P3 p[4], int lazy, double out[3]) {
    const Buf b{circumcenter_workspace(), 1};
    double o[3] = {0.0, 0.0, 0.0};
    int need = 7;
    if (lazy) need = circ_lazy_pass(p[0], p[1], p[2], p[3], o);
    if (need) {
        // the small layout first, the worst-case one if a step outgrows it,
        // as the GPU's two exact passes do
        bool over = false;
        P3 e = circ_exact(p[0], p[1], p[2], p[3], b, kCircSmall, &over);
        if (over) e = circ_exact(p[0], p[1], p[2], p[3], b, kCircFull, &over);
        if (need & 1) o[0] = e.x;
        if (need & 2) o[1] = e.y;
        if (need & 4) o[2] = e.z;
    }
    for (int k = 0; k < 3; ++k) out[k] = o[k];
}

}  // namespace d3
