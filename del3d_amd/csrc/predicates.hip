// predicates.hip - batched exact orient3d / in_sphere for gfx950.
//
// Three kernels per batch, stream-ordered, no host round trip in between:
//
//  pred_fast   one lane per query. Index records are staged through LDS with
//              fully coalesced dword loads (a 256-query tile is 5 KB of
//              indices), the tile is regrouped so that in_sphere and orient3d
//              queries occupy separate wavefronts, then each lane gathers its
//              4-5 points and runs the FP64 filter in registers. Undecided
//              queries try the certified small-grid tier (T1) and, for
//              exactly cospherical in_sphere queries, the whole symbolic
//              perturbation, still in registers. Whatever is left is appended
//              (one atomic per wavefront) to a compacted queue.
//  pred_exact  persistent grid over the queue: exact integer arithmetic with
//              1-2 limbs per difference (T2), perturbation included. Queries
//              whose exponent span needs more limbs go to a second queue.
//  pred_wide   one query per workgroup, exact integer arithmetic of any width
//              in LDS (T3).
//
// All device code is compiled with -ffp-contract=off: the filter must round
// exactly as the reference's (src/predicates.cpp:125-201) does.
#include "common.h"
#include "pred_core.h"

namespace d3 {

constexpr int kTile = 256;           // queries per fast-kernel tile (one workgroup)
constexpr int kWideLimbsMax = 67;    // > limbs_for() of any finite input (66)
constexpr int kWideBlock = 64;

struct PredCounters {
    uint32_t queued;  // entries in the exact queue
    uint32_t wide;    // entries in the wide queue
};

__device__ inline uint64_t lanemask_lt() { return __lanemask_lt(); }

// Appends `q` for every lane with `want`; one atomic per wavefront.
__device__ inline void wave_push(bool want, uint32_t q, uint32_t* counter, uint32_t* list) {
    const uint64_t m = __ballot(want);
    if (m == 0) return;
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(m));
    base = __shfl(base, leader);
    if (want) list[base + (uint32_t)__popcll(m & lanemask_lt())] = q;
}

__device__ inline void wave_count(bool flag, unsigned long long* slot) {
    const uint64_t m = __ballot(flag);
    if (m && __lane_id() == __ffsll((long long)m) - 1) atomicAdd(slot, (unsigned long long)__popcll(m));
}

// ---------------------------------------------------------------------------
// Fast path
// ---------------------------------------------------------------------------
template <bool kStats>
__global__ __launch_bounds__(kTile) void pred_fast(const double* __restrict__ xyz,
                                                   const uint8_t* __restrict__ op,
                                                   const int32_t* __restrict__ idx, uint32_t n_q,
                                                   int8_t* __restrict__ out,
                                                   uint32_t* __restrict__ queue,
                                                   PredCounters* __restrict__ ctr,
                                                   unsigned long long* __restrict__ stats) {
    __shared__ int32_t s_idx[kTile * 5];
    __shared__ uint8_t s_op[kTile];
    __shared__ uint16_t s_perm[kTile];
    __shared__ uint32_t s_wave_ins[kTile / 64];

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = __lane_id();

    for (uint32_t base = blockIdx.x * kTile; base < n_q; base += gridDim.x * kTile) {
        const uint32_t cnt = min((uint32_t)kTile, n_q - base);
        __syncthreads();  // the previous tile is done with the LDS arrays
        // Stage the tile's index records: 5*cnt consecutive int32, one dword
        // per lane per step, consecutive lanes on consecutive addresses.
        const int32_t* src = idx + (size_t)base * 5;
        for (int k = tid; k < (int)cnt * 5; k += kTile) s_idx[k] = src[k];
        const bool live = (uint32_t)tid < cnt;
        const uint8_t my_op = live ? op[base + tid] : (uint8_t)DEL3D_OP_ORIENT3D;

        // Regroup: in_sphere-type queries first, orient3d after, each in
        // query order, so wavefronts mostly run one code path.
        const bool is_ins = live && my_op != DEL3D_OP_ORIENT3D;
        const uint64_t ins_mask = __ballot(is_ins);
        if (lane == 0) s_wave_ins[wave] = (uint32_t)__popcll(ins_mask);
        __syncthreads();
        uint32_t ins_before = 0, ins_total = 0;
#pragma unroll
        for (int w = 0; w < kTile / 64; ++w) {
            const uint32_t c = s_wave_ins[w];
            ins_total += c;
            ins_before += (w < wave) ? c : 0u;
        }
        const uint32_t lt = (uint32_t)__popcll(ins_mask & lanemask_lt());
        // live orient3d lanes before this one = live lanes before - in_sphere lanes before
        const uint32_t pos = is_ins ? ins_before + lt
                                    : ins_total + ((uint32_t)(wave * 64 + lane) - (ins_before + lt));
        if (live) s_perm[pos] = (uint16_t)tid;
        s_op[tid] = my_op;
        __syncthreads();

        const bool act = (uint32_t)tid < cnt;
        const int lq = act ? s_perm[tid] : 0;
        const uint32_t q = base + lq;
        const uint8_t o = s_op[lq];
        int32_t id[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) id[k] = s_idx[lq * 5 + k];

        int sign = 0;
        bool done = false;
        int tier = kTierFilter;
        bool perturbed = false;
        if (act) {
            P3 p[5];
            p[0] = load_p3(xyz, id[0]);
            p[1] = load_p3(xyz, id[1]);
            p[2] = load_p3(xyz, id[2]);
            p[3] = load_p3(xyz, id[3]);
            if (o == DEL3D_OP_ORIENT3D) {
                const Filt f = orient_filter(p[0], p[1], p[2], p[3]);
                if (filter_decides(f, kOrientBound)) {
                    sign = -sgn(f.det);
                    done = true;
                } else if (t1_certify(p, 4, f.maxd, 15, -358)) {
                    sign = -sgn(f.det);
                    done = true;
                    tier = kTierCert;
                }
            } else {
                p[4] = load_p3(xyz, id[4]);
                const Filt f = insphere_filter(p[0], p[1], p[2], p[3], p[4]);
                if (filter_decides(f, kInSphereBound)) {
                    sign = -sgn(f.det);
                    done = true;
                } else if (t1_certify(p, 5, f.maxd, 8, -214)) {
                    tier = kTierCert;
                    sign = -sgn(f.det);
                    done = true;
                    if (sign == 0 && o == DEL3D_OP_IN_SPHERE) {
                        perturbed = true;
                        // every four of these points certify at degree 3 too,
                        // so the FP64 orient3d determinant is exact here
                        sign = perturbation_walk(p, id, [&](int k) {
                            P3 t[4];
                            subst_tet(p, k, t);
                            return -sgn(orient_filter(t[0], t[1], t[2], t[3]).det);
                        });
                    }
                }
            }
            if (done) out[q] = (int8_t)sign;
        }
        wave_push(act && !done, q, &ctr->queued, queue);
        if (kStats) {
            wave_count(act && done && tier == kTierFilter, &stats[0]);
            wave_count(act && done && tier == kTierCert, &stats[1]);
            wave_count(act && perturbed, &stats[4]);
        }
    }
}

// ---------------------------------------------------------------------------
// Exact path, registers (1..3 limbs)
// ---------------------------------------------------------------------------
template <int L>
__device__ int exact_query(uint8_t o, const P3 p[5], const int32_t id[5], int q, bool& perturbed) {
    if (o == DEL3D_OP_ORIENT3D) {
        return -int_orient_raw<L>(to_lp3<L>(p[0], q), to_lp3<L>(p[1], q), to_lp3<L>(p[2], q),
                                  to_lp3<L>(p[3], q));
    }
    const int raw = -int_insphere_raw<L>(to_lp3<L>(p[0], q), to_lp3<L>(p[1], q),
                                         to_lp3<L>(p[2], q), to_lp3<L>(p[3], q),
                                         to_lp3<L>(p[4], q));
    if (raw != 0 || o == DEL3D_OP_IN_SPHERE_UNPERTURBED) return raw;
    perturbed = true;
    return perturbation_walk(p, id, [&](int k) {
        P3 t[4];
        subst_tet(p, k, t);
        const Filt f = orient_filter(t[0], t[1], t[2], t[3]);
        if (filter_decides(f, kOrientBound)) return -sgn(f.det);
        return orient_public_int<L>(t, q);
    });
}

__global__ __launch_bounds__(256) void pred_exact(const double* __restrict__ xyz,
                                                  const uint8_t* __restrict__ op,
                                                  const int32_t* __restrict__ idx,
                                                  int8_t* __restrict__ out,
                                                  const uint32_t* __restrict__ queue,
                                                  uint32_t* __restrict__ wide,
                                                  PredCounters* __restrict__ ctr,
                                                  unsigned long long* __restrict__ stats) {
    const uint32_t n = ctr->queued;
    const uint32_t stride = gridDim.x * blockDim.x;
    // i0 is the wavefront's base, so the trip count is uniform per wavefront
    // and the ballots in wave_count see every lane
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        const bool act = i < n;
        bool perturbed = false, done = false;
        if (act) {
            const uint32_t q = queue[i];
            const uint8_t o = op[q];
            int32_t id[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) id[k] = idx[(size_t)q * 5 + k];
            P3 p[5];
            const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
            for (int k = 0; k < np; ++k) p[k] = load_p3(xyz, id[k]);
            if (np == 4) p[4] = p[3];
            const Grid g = grid_of(p, np);
            const int L = limbs_for(g);
            int s = 0;
            done = true;
            if (L == 1) {
                s = exact_query<1>(o, p, id, g.q, perturbed);
            } else if (L == 2) {
                s = exact_query<2>(o, p, id, g.q, perturbed);
            } else {
                done = false;
                wide[atomicAdd(&ctr->wide, 1u)] = q;  // rare: exponent span above 2^55
            }
            if (done) out[q] = (int8_t)s;
        }
        if (stats) {
            wave_count(act && done, &stats[2]);
            wave_count(act && perturbed, &stats[4]);
        }
    }
}

// ---------------------------------------------------------------------------
// Exact path, any width, LDS (one query per workgroup, lane 0 computes)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kWideBlock) void pred_wide(const double* __restrict__ xyz,
                                                        const uint8_t* __restrict__ op,
                                                        const int32_t* __restrict__ idx,
                                                        int8_t* __restrict__ out,
                                                        const uint32_t* __restrict__ wide,
                                                        const PredCounters* __restrict__ ctr,
                                                        unsigned long long* __restrict__ stats) {
    __shared__ uint32_t ws[64 * kWideLimbsMax];
    const uint32_t n = ctr->wide;
    if (threadIdx.x != 0) return;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        const uint32_t q = wide[i];
        const uint8_t o = op[q];
        int32_t id[5];
        for (int k = 0; k < 5; ++k) id[k] = idx[(size_t)q * 5 + k];
        P3 p[5];
        const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
        for (int k = 0; k < np; ++k) p[k] = load_p3(xyz, id[k]);
        if (np == 4) p[4] = p[3];
        const Grid g = grid_of(p, np);
        const int L = limbs_for(g);
        int s;
        if (o == DEL3D_OP_ORIENT3D) {
            s = -rt_orient_raw(p[0], p[1], p[2], p[3], g.q, L, ws);
        } else {
            s = -rt_insphere_raw(p[0], p[1], p[2], p[3], p[4], g.q, L, ws);
            if (s == 0 && o == DEL3D_OP_IN_SPHERE) {
                if (stats) atomicAdd(&stats[4], 1ull);
                s = perturbation_walk(p, id, [&](int k) {
                    P3 t[4];
                    subst_tet(p, k, t);
                    const Filt f = orient_filter(t[0], t[1], t[2], t[3]);
                    if (filter_decides(f, kOrientBound)) return -sgn(f.det);
                    return -rt_orient_raw(t[0], t[1], t[2], t[3], g.q, L, ws);
                });
            }
        }
        out[q] = (int8_t)s;
        if (stats) atomicAdd(&stats[3], 1ull);
    }
}

// ---------------------------------------------------------------------------
// Host launcher
// ---------------------------------------------------------------------------
constexpr uint32_t kChunk = 1u << 26;  // queries per launch sequence (bounds the queues)

// When `ev` is given (6 events), the kernels of the (single) chunk are bracketed
// by ev[0..1] (fast), ev[2..3] (exact) and ev[4..5] (wide).
void launch_predicates(const double* xyz, const uint8_t* op, const int32_t* idx, int64_t n_q,
                       int8_t* out, hipStream_t s, unsigned long long* stats_dev,
                       hipEvent_t* ev) {
    if (n_q <= 0) return;
    const uint32_t chunk = (uint32_t)std::min<int64_t>(n_q, kChunk);
    Scratch sc(s);
    uint32_t* queue = sc.get<uint32_t>(chunk);
    uint32_t* wide = sc.get<uint32_t>(chunk);
    PredCounters* ctr = sc.get<PredCounters>(1);
    const int cus = device_cu_count();
    for (int64_t off = 0; off < n_q; off += chunk) {
        const uint32_t n = (uint32_t)std::min<int64_t>(chunk, n_q - off);
        fill_u32(reinterpret_cast<uint32_t*>(ctr), 0u, sizeof(PredCounters) / 4, s);
        const uint32_t tiles = (n + kTile - 1) / kTile;
        const uint32_t fast_grid = std::min<uint32_t>(tiles, (uint32_t)cus * 16);
        const uint8_t* opc = op + off;
        const int32_t* idc = idx + off * 5;
        int8_t* outc = out + off;
        if (ev) D3_HIP(hipEventRecord(ev[0], s));
        if (stats_dev) {
            pred_fast<true><<<fast_grid, kTile, 0, s>>>(xyz, opc, idc, n, outc, queue,
                                                        ctr, stats_dev);
            D3_LAUNCH_CHECK();
        } else {
            pred_fast<false><<<fast_grid, kTile, 0, s>>>(xyz, opc, idc, n, outc,
                                                         queue, ctr, nullptr);
            D3_LAUNCH_CHECK();
        }
        if (ev) {
            D3_HIP(hipEventRecord(ev[1], s));
            D3_HIP(hipEventRecord(ev[2], s));
        }
        pred_exact<<<cus * 4, 256, 0, s>>>(xyz, opc, idc, outc, queue,
                                           wide, ctr, stats_dev);
        D3_LAUNCH_CHECK();
        if (ev) {
            D3_HIP(hipEventRecord(ev[3], s));
            D3_HIP(hipEventRecord(ev[4], s));
        }
        pred_wide<<<cus * 4, kWideBlock, 0, s>>>(xyz, opc, idc, outc, wide, ctr,
                                                 stats_dev);
        D3_LAUNCH_CHECK();
        if (ev) D3_HIP(hipEventRecord(ev[5], s));
    }
}

}  // namespace d3

namespace d3 {

// ---------------------------------------------------------------------------
// One query, every tier in one thread: the scalar entry points (C++ drop-in).
// ---------------------------------------------------------------------------
struct ScalarQuery {
    double xyz[15];
    int32_t id[5];
    uint8_t op;
    int8_t sign;
};

__global__ __launch_bounds__(64) void pred_single(ScalarQuery* qd) {
    __shared__ uint32_t ws[64 * kWideLimbsMax];
    if (threadIdx.x != 0) return;
    const uint8_t o = qd->op;
    const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
    P3 p[5];
    int32_t id[5];
    for (int k = 0; k < 5; ++k) id[k] = qd->id[k];
    // slots hold their own coordinates: xyz[3k..3k+2] is slot k's point, and
    // id[] carries the caller's pointer identities
    for (int k = 0; k < np; ++k) p[k] = P3{qd->xyz[3 * k], qd->xyz[3 * k + 1], qd->xyz[3 * k + 2]};
    if (np == 4) p[4] = p[3];
    const Grid g = grid_of(p, np);
    const int L = limbs_for(g);
    auto orient_pub = [&](const P3* t) -> int {
        const Filt f = orient_filter(t[0], t[1], t[2], t[3]);
        if (filter_decides(f, kOrientBound)) return -sgn(f.det);
        if (L == 1) return orient_public_int<1>(t, g.q);
        if (L == 2) return orient_public_int<2>(t, g.q);
        return -rt_orient_raw(t[0], t[1], t[2], t[3], g.q, L, ws);
    };
    int s;
    if (o == DEL3D_OP_ORIENT3D) {
        s = orient_pub(p);
    } else {
        const Filt f = insphere_filter(p[0], p[1], p[2], p[3], p[4]);
        if (filter_decides(f, kInSphereBound)) {
            s = -sgn(f.det);
        } else if (L <= 2) {
            bool dummy = false;
            s = L == 1 ? exact_query<1>(DEL3D_OP_IN_SPHERE_UNPERTURBED, p, id, g.q, dummy)
                       : exact_query<2>(DEL3D_OP_IN_SPHERE_UNPERTURBED, p, id, g.q, dummy);
        } else {
            s = -rt_insphere_raw(p[0], p[1], p[2], p[3], p[4], g.q, L, ws);
        }
        if (s == 0 && o == DEL3D_OP_IN_SPHERE) {
            s = perturbation_walk(p, id, [&](int k) {
                P3 t[4];
                subst_tet(p, k, t);
                return orient_pub(t);
            });
        }
    }
    qd->sign = (int8_t)s;
}

void launch_single(ScalarQuery* qd, hipStream_t s) {
    pred_single<<<1, 64, 0, s>>>(qd);
    D3_LAUNCH_CHECK();
}

}  // namespace d3
