// circumcenter.hip - batched circumcentres of tetrahedra for gfx950.
//
// Replaces, per cell, one call of the reference's circumcenter or
// circumcenter_lazy (src/triangulation.cpp:805-930; include/del3d/delaunay.h:
// 29-49). Two kernels per batch on the caller's stream:
//
//  circ_lazy_k   one lane per cell: the interval pass under (emulated) upward
//                rounding; accepted coordinates are written, a cell with a
//                coordinate left to the exact form is appended, one atomic
//                per wavefront, to a queue with its 3-bit coordinate mask;
//  circ_exact_k  one lane per queued cell (or per cell, for the exact form):
//                the exact closed form in expansion arithmetic on a strided
//                per-lane workspace in device memory, first with a small
//                workspace that typical cells fit, then, for the cells whose
//                expansions outgrow it, with the worst-case workspace.
//
// Cells with an index outside [0, n_pts) get NaN coordinates.
#include "circ_core.h"
#include "common.h"
#include "predicates.h"

namespace d3 {

// Lanes of the exact passes: the small-workspace pass (11.6 KB per lane, so
// 1.5 GB of scratch at most; 131072 lanes ran 2.6x faster than 32768 and
// 262144 no faster) and the worst-case pass for the cells that overflow it
// (79 KB per lane).
constexpr int kCircSmallLanes = 131072;
constexpr int kCircFullLanes = 1024;

__device__ inline bool load_cell(const double* xyz, int64_t n_pts, const int32_t* cells, uint32_t c,
                                 P3 p[4]) {
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int32_t i = cells[(size_t)c * 4 + k];
        ok = ok && (uint32_t)i < (uint64_t)n_pts;
        p[k] = ok ? load_p3(xyz, i) : P3{0.0, 0.0, 0.0};
    }
    return ok;
}

__global__ __launch_bounds__(256) void circ_lazy_k(const double* __restrict__ xyz, int64_t n_pts,
                                                   const int32_t* __restrict__ cells, uint32_t n,
                                                   double* __restrict__ out, uint32_t* __restrict__ queue,
                                                   uint8_t* __restrict__ need, uint32_t* __restrict__ counter) {
    for (uint32_t c0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); c0 < n; c0 += gridDim.x * blockDim.x) {
        const uint32_t c = c0 + __lane_id();
        int mask = 0;
        if (c < n) {
            P3 p[4];
            double o[3] = {__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
            if (load_cell(xyz, n_pts, cells, c, p)) mask = circ_lazy_pass(p[0], p[1], p[2], p[3], o);
#pragma unroll
            for (int k = 0; k < 3; ++k)
                if (!(mask >> k & 1)) out[(size_t)c * 3 + k] = o[k];
        }
        const bool push = mask != 0;
        const uint64_t m = __ballot(push);
        if (m) {
            const int leader = __ffsll((long long)m) - 1;
            uint32_t base = 0;
            if (__lane_id() == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
            base = __shfl(base, leader);
            if (push) {
                const uint32_t at = base + (uint32_t)__popcll(m & __lanemask_lt());
                queue[at] = c;
                need[at] = (uint8_t)mask;
            }
        }
    }
}

// One exact pass. all_cells: every cell, all three coordinates (the exact
// form); else the queued cells and their masks. kSmall: the small workspace
// layout, and a cell that overflows it goes to the retry queue (with its
// mask) for the full-layout pass.
template <bool kSmall>
__global__ __launch_bounds__(256) void circ_exact_k(const double* __restrict__ xyz, int64_t n_pts,
                                                    const int32_t* __restrict__ cells, uint32_t n_cells,
                                                    double* __restrict__ out,
                                                    const uint32_t* __restrict__ queue,
                                                    const uint8_t* __restrict__ need,
                                                    const uint32_t* __restrict__ counter, int all_cells,
                                                    double* __restrict__ ws_base,
                                                    uint32_t* __restrict__ retry,
                                                    uint8_t* __restrict__ retry_need,
                                                    uint32_t* __restrict__ retry_counter) {
    const uint32_t lanes = gridDim.x * blockDim.x;
    const uint32_t lane = blockIdx.x * blockDim.x + threadIdx.x;
    const Buf ws{ws_base + lane, (int64_t)lanes};
    const CircLayout L = kSmall ? kCircSmall : kCircFull;
    const uint32_t n = all_cells ? n_cells : *counter;
    // wave-uniform trip count: every lane of a wavefront runs the same
    // iterations, so the retry push below is reached convergently
    for (uint32_t i0 = lane - __lane_id(); i0 < n; i0 += lanes) {
        const uint32_t i = i0 + __lane_id();
        bool over = false;
        uint32_t c = 0;
        int mask = 0;
        if (i < n) {
            c = all_cells ? i : queue[i];
            mask = all_cells ? 7 : need[i];
            P3 p[4];
            P3 r{__builtin_nan(""), __builtin_nan(""), __builtin_nan("")};
            if (load_cell(xyz, n_pts, cells, c, p)) r = circ_exact(p[0], p[1], p[2], p[3], ws, L, &over);
            if (!over) {
                if (mask & 1) out[(size_t)c * 3 + 0] = r.x;
                if (mask & 2) out[(size_t)c * 3 + 1] = r.y;
                if (mask & 4) out[(size_t)c * 3 + 2] = r.z;
            }
        }
        if (kSmall) {
            const uint64_t m = __ballot(over);
            if (m) {
                const int leader = __ffsll((long long)m) - 1;
                uint32_t base = 0;
                if (__lane_id() == leader) base = atomicAdd(retry_counter, (uint32_t)__popcll(m));
                base = __shfl(base, leader);
                if (over) {
                    const uint32_t at = base + (uint32_t)__popcll(m & __lanemask_lt());
                    retry[at] = c;
                    retry_need[at] = (uint8_t)mask;
                }
            }
        }
    }
}

__global__ void k_clear_two(uint32_t* c) { c[0] = 0u; c[1] = 0u; }

void launch_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n_cells,
                          int lazy, double* out, hipStream_t s) {
    if (n_cells <= 0) return;
    const uint32_t chunk = (uint32_t)std::min<int64_t>(n_cells, kChunk);
    const uint32_t lanes_s = (uint32_t)std::min<int64_t>(((n_cells + 255) / 256) * 256, kCircSmallLanes);
    const uint32_t lanes_f = (uint32_t)std::min<int64_t>(((n_cells + 255) / 256) * 256, kCircFullLanes);
    Scratch sc(s);
    double* ws_s = sc.get<double>((size_t)lanes_s * kCircSmall.total());
    double* ws_f = sc.get<double>((size_t)lanes_f * kCircFull.total());
    uint32_t* queue = lazy ? sc.get<uint32_t>(chunk) : nullptr;
    uint8_t* need = lazy ? sc.get<uint8_t>(chunk) : nullptr;
    uint32_t* retry = sc.get<uint32_t>(chunk);
    uint8_t* retry_need = sc.get<uint8_t>(chunk);
    uint32_t* counters = sc.get<uint32_t>(2);  // [0] lazy queue, [1] retry queue
    const int cus = device_cu_count();
    for (int64_t off = 0; off < n_cells; off += chunk) {
        const uint32_t n = (uint32_t)std::min<int64_t>(chunk, n_cells - off);
        k_clear_two<<<1, 1, 0, s>>>(counters);
        D3_LAUNCH_CHECK();
        if (lazy) {
            const unsigned g = (unsigned)std::min<uint32_t>((n + 255) / 256, (uint32_t)cus * 16);
            circ_lazy_k<<<g, 256, 0, s>>>(xyz, n_pts, cells + off * 4, n, out + off * 3, queue, need,
                                          counters);
            D3_LAUNCH_CHECK();
        }
        circ_exact_k<true><<<lanes_s / 256, 256, 0, s>>>(xyz, n_pts, cells + off * 4, n, out + off * 3,
                                                         queue, need, counters, lazy ? 0 : 1, ws_s, retry,
                                                         retry_need, counters + 1);
        D3_LAUNCH_CHECK();
        circ_exact_k<false><<<lanes_f / 256, 256, 0, s>>>(xyz, n_pts, cells + off * 4, n, out + off * 3,
                                                          retry, retry_need, counters + 1, 0, ws_f,
                                                          nullptr, nullptr, nullptr);
        D3_LAUNCH_CHECK();
    }
}

}  // namespace d3
