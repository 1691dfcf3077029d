// predicates2.hip - batched coplanar predicates for gfx950.
//
// Replaces, per query, one call of the reference's coplanar predicates
// (src/predicates2.cpp:111-259; include/del3d/predicates2.h:26-69) or of the
// Delaunay driver's plain-double orientation (src/triangulation.cpp:296-308). Two
// kernels per batch on the caller's stream:
//
//  pred2_fast   one lane per query: the whole predicate (perturbation
//               included) with the filters in the reference's operation order
//               and a certified FP64 evaluation for points on a small
//               power-of-two grid; a query any step leaves undecided is
//               appended, one atomic per wavefront, to a queue;
//  pred2_exact  the whole predicate with exact integer fallbacks and the
//               cocircular perturbation: first one lane per query on a
//               private device-memory workspace for queries of up to 4 limbs
//               (ordinary doubles), then one query per workgroup on a
//               workspace in LDS sized for any finite input.
//
// The coplanar predicates are called rarely by the reference's driver (during
// the dimension bootstrap and on exactly coplanar hull facets).
#include "common.h"
#include "pred2_core.h"
#include "predicates.h"

namespace d3 {

constexpr int kWs2 = 207 * kMaxLimbs2 + 16;  // in_circle_workspace_limbs(kMaxLimbs2)

// Queries whose op byte is below op_base belong to another kernel of a mixed
// batch and are left untouched; the others have op code (byte - op_base).
__global__ __launch_bounds__(256) void pred2_fast_k(const double* __restrict__ xyz, int64_t n_pts,
                                                    const uint8_t* __restrict__ op,
                                                    const int32_t* __restrict__ idx, uint32_t n_q,
                                                    int8_t* __restrict__ out, uint32_t* __restrict__ queue,
                                                    uint32_t* __restrict__ counter, uint8_t op_base) {
    for (uint32_t q0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); q0 < n_q;
         q0 += gridDim.x * blockDim.x) {
        const uint32_t q = q0 + __lane_id();
        bool pending = false;
        if (q < n_q && op[q] >= op_base) {
            const uint8_t o = (uint8_t)(op[q] - op_base);
            const int np = points_of_op2(o);
            P3 pt[4];
            int32_t id[4];
            bool valid = o <= kOp2Last;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                id[k] = idx[(size_t)q * 5 + k];
                const bool used = k < np;
                valid = valid && (!used || (uint32_t)id[k] < (uint64_t)n_pts);
                pt[k] = (used && valid) ? load_p3(xyz, id[k]) : P3{0.0, 0.0, 0.0};
            }
            int s = 0;
            if (valid) {
                s = pred2_fast(o, pt, id);
                pending = s == kPending2;
            }
            if (!pending) out[q] = (int8_t)s;
        }
        const uint64_t m = __ballot(pending);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (pending) queue[base + (uint32_t)__popcll(m & __lanemask_lt())] = q;
        }
    }
}

// Exact pass 1: one lane per queued query whose points need at most
// kLaneLimbs limbs (every query of ordinary doubles), on a private workspace
// in device memory; wider queries go on to the retry queue.
constexpr int kLaneLimbs = 4;
constexpr int kLaneWs = 207 * kLaneLimbs + 16;  // in_circle_workspace_limbs(kLaneLimbs)
constexpr int kLaneExactLanes = 65536;           // 65536 x 3.4 KB of workspace

__global__ __launch_bounds__(256) void pred2_exact_lanes_k(const double* __restrict__ xyz,
                                                           const uint8_t* __restrict__ op,
                                                           const int32_t* __restrict__ idx,
                                                           int8_t* __restrict__ out,
                                                           const uint32_t* __restrict__ queue,
                                                           const uint32_t* __restrict__ counter,
                                                           uint8_t op_base, uint32_t* __restrict__ ws_base,
                                                           uint32_t* __restrict__ retry,
                                                           uint32_t* __restrict__ retry_counter) {
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t* ws = ws_base + (size_t)lane * kLaneWs;
    const uint32_t n = *counter;
    // wave-uniform trip count, so the retry push is reached convergently
    for (uint32_t i0 = lane - __lane_id(); i0 < n; i0 += lanes) {
        const uint32_t i = i0 + __lane_id();
        bool wide = false;
        uint32_t q = 0;
        if (i < n) {
            q = queue[i];
            const uint8_t o = (uint8_t)(op[q] - op_base);
            const int np = points_of_op2(o);
            P3 pt[4];
            int32_t id[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                id[k] = idx[(size_t)q * 5 + k];
                pt[k] = k < np ? load_p3(xyz, id[k]) : P3{0.0, 0.0, 0.0};
            }
            // the limbs of every exact evaluation of this query are bounded
            // by those of all its points together
            wide = limbs_for(grid_of(pt, np)) > kLaneLimbs;
            if (!wide) out[q] = (int8_t)pred2_eval(o, pt, id, ws);
        }
        const uint64_t m = __ballot(wide);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (__lane_id() == leader) base = atomicAdd(retry_counter, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (wide) retry[base + (uint32_t)__popcll(m & __lanemask_lt())] = q;
        }
    }
}

// Exact pass 2: one query per workgroup, any width, in LDS.
__global__ __launch_bounds__(64) void pred2_exact_k(const double* __restrict__ xyz,
                                                    const uint8_t* __restrict__ op,
                                                    const int32_t* __restrict__ idx,
                                                    int8_t* __restrict__ out,
                                                    const uint32_t* __restrict__ queue,
                                                    const uint32_t* __restrict__ counter,
                                                    uint8_t op_base) {
    __shared__ uint32_t ws[kWs2];
    if (threadIdx.x != 0) return;
    const uint32_t n = *counter;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t q = queue[i];
        const uint8_t o = (uint8_t)(op[q] - op_base);
        const int np = points_of_op2(o);
        P3 pt[4];
        int32_t id[4];
        for (int k = 0; k < 4; ++k) {
            id[k] = idx[(size_t)q * 5 + k];
            pt[k] = k < np ? load_p3(xyz, id[k]) : P3{0.0, 0.0, 0.0};
        }
        out[q] = (int8_t)pred2_eval(o, pt, id, ws);
    }
}

__global__ void k_clear_u32x2(uint32_t* c) {
    c[0] = 0u;
    c[1] = 0u;
}

void launch_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* out, hipStream_t s, uint8_t op_base) {
    if (n_q <= 0) return;
    const uint32_t chunk = (uint32_t)std::min<int64_t>(n_q, kChunk);
    Scratch sc(s);
    uint32_t* queue = sc.get<uint32_t>(chunk);
    uint32_t* retry = sc.get<uint32_t>(chunk);
    uint32_t* counters = sc.get<uint32_t>(2);  // [0] queue, [1] retry queue
    // lanes of the first exact pass (at most the batch's size, so a small
    // batch takes a small workspace)
    const uint32_t lanes = (uint32_t)std::min<int64_t>(((n_q + 255) / 256) * 256, kLaneExactLanes);
    uint32_t* ws = sc.get<uint32_t>((size_t)lanes * kLaneWs);
    const int cus = device_cu_count();
    for (int64_t off = 0; off < n_q; off += chunk) {
        const uint32_t n = (uint32_t)std::min<int64_t>(chunk, n_q - off);
        k_clear_u32x2<<<1, 1, 0, s>>>(counters);
        D3_LAUNCH_CHECK();
        const unsigned g = (unsigned)std::min<uint32_t>((n + 255) / 256, (uint32_t)cus * 16);
        pred2_fast_k<<<g, 256, 0, s>>>(xyz, n_pts, op + off, idx + off * 5, n, out + off, queue, counters,
                                       op_base);
        D3_LAUNCH_CHECK();
        pred2_exact_lanes_k<<<lanes / 256, 256, 0, s>>>(xyz, op + off, idx + off * 5, out + off, queue,
                                                        counters, op_base, ws, retry, counters + 1);
        D3_LAUNCH_CHECK();
        pred2_exact_k<<<cus * 2, 64, 0, s>>>(xyz, op + off, idx + off * 5, out + off, retry, counters + 1,
                                             op_base);
        D3_LAUNCH_CHECK();
    }
}

}  // namespace d3
