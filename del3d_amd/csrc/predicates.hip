// predicates.hip - batched exact orient3d / in_sphere for gfx950.
//
// A batch is a list of tasks, each a contiguous range of queries of one point
// set (a plain call is one set). Kernels, stream-ordered, no host round trip:
//
//  pred_batch_lds     sets of at most kLdsPts points: one 1024-thread
//                     workgroup per CU holds the whole set in LDS (up to
//                     132 KB); its 16 wavefronts then take chunks of the
//                     task's queries independently, each regrouping its
//                     chunk through its own table in LDS.
//  pred_batch_packed  single sets of kLdsPts+1 .. kPackPts points on a small
//                     integer grid (pack_point): the same, with the set's
//                     packed copy (4 bytes a point) in LDS and every query
//                     decided on the grid (packed_decide).
//  pred_batch_global  larger sets: 256-thread workgroups gather points from
//                     global memory (L2 / Infinity Cache / HBM), with each
//                     256-query tile's index records staged through LDS with
//                     coalesced dword loads and regrouped across the
//                     workgroup; for a single packable set,
//                     pred_batch_global_packed gathers its 4-byte words and
//                     decides on the grid.
//     All: in_sphere and orient3d queries occupy separate wavefronts as far
//     as possible, and each lane runs, in registers, the FP64 filter in the
//     reference's operation order, then for undecided queries the certified
//     small-grid tier (T1) and, for exactly cospherical in_sphere queries,
//     the whole symbolic perturbation. Whatever is left goes to a compacted
//     queue, through a per-wavefront buffer in LDS (WaveQueue).
//  pred_exact         persistent grid over the queue: exact integer arithmetic
//                     with 1-2 limbs per difference (T2), perturbation
//                     included; wider exponent spans go to a second queue.
//  pred_exact_mid     the wide queue's 3-4 limb queries, in registers.
//  pred_wide          one query per workgroup, exact integers of any width in
//                     LDS, the products of each level on separate lanes (T3).
//
// Why LDS for small sets: a query gathers 4-5 random points of 24 bytes, two
// L2 requests each (a 16- and an 8-byte load), so with points in L2 the filter
// is bound by the L2 request rate (16 requests per clock per XCD), not by
// FP64. From LDS the same gathers cost a few bank cycles.
//
// All device code is compiled with -ffp-contract=off: the filter must round
// exactly as the reference's (src/predicates.cpp:125-201) does.
#include <algorithm>

#include <mutex>

#include "common.h"
#include "pred_core.h"
#include "predicates.h"

namespace d3 {

constexpr int kWideLimbsMax = 67;  // > limbs_for() of any finite input (66)
constexpr int kWideBlock = 64;

// Out-of-range point indices are undefined behaviour in the reference (a bad
// pointer); here such a query, or one with an unknown op code, gets sign 0 and
// is counted in stats[5], so a bad index can never make the device read
// outside the point array.
__device__ inline uint64_t lanemask_lt() { return __lanemask_lt(); }

__device__ inline void wave_count(bool flag, unsigned long long* slot) {
    const uint64_t m = __ballot(flag);
    if (m && __lane_id() == __ffsll((long long)m) - 1) atomicAdd(slot, (unsigned long long)__popcll(m));
}

// Orders this wavefront's LDS accesses (a buffer written by some lanes and
// read by others) without a workgroup barrier: a wavefront's LDS
// instructions execute in order, so this only stops the compiler from
// moving them.
__device__ inline void wave_lds_order() { __builtin_amdgcn_wave_barrier(); }

// A wavefront's entries for the exact queue, gathered in its own LDS buffer
// of kCap entries and appended to the queue in blocks of at least kCap/2:
// an atomic on the queue's counter costs about as much as a query's whole
// filter when every wavefront of the device adds to that one address
// (device-wide atomics are done in memory, one at a time), so one per block
// instead of one per wavefront step. flush() before the wavefront ends.
template <int kCap>
struct WaveQueue {
    uint32_t* buf;  // kCap entries in LDS
    uint32_t n;     // entries held (the same in every lane)
    __device__ void flush(uint32_t* counter, uint32_t* list) {
        if (n == 0) return;
        wave_lds_order();
        const int lane = __lane_id();
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(counter, n);
        base = __builtin_amdgcn_readfirstlane(__shfl(base, 0));
        for (uint32_t i = (uint32_t)lane; i < n; i += 64) list[base + i] = buf[i];
        wave_lds_order();  // the reads come before the next entries' writes
        n = 0;
    }
    __device__ void push(bool want, uint32_t q, uint32_t* counter, uint32_t* list) {
        const uint64_t m = __ballot(want);
        if (m == 0) return;
        if (want) buf[n + (uint32_t)__popcll(m & lanemask_lt())] = q;
        n += (uint32_t)__popcll(m);  // n < kCap/2 before, so at most kCap/2 + 64
        if (n >= (uint32_t)kCap / 2) flush(counter, list);
    }
};

// A wavefront's statistics counts (stats slots 0, 1, 4, 5), added to the
// device counters once, by flush().
struct WaveStats {
    uint64_t c[4] = {0, 0, 0, 0};
    __device__ void count(int k, bool flag) { c[k] += (uint64_t)__popcll(__ballot(flag)); }
    __device__ void flush(unsigned long long* stats) {
        const int slot[4] = {0, 1, 4, 5};
        if (__lane_id() == 0) {
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (c[k]) atomicAdd(&stats[slot[k]], (unsigned long long)c[k]);
        }
    }
};

// A value every lane holds equally, moved to scalar registers.
__device__ inline int64_t uniform64(int64_t v) {
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)(uint64_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// Point fetch from global memory (set base already applied) or from LDS.
// use32 (uniform per task): the set's float copy, one 16-byte load per point
// instead of a 16- and an 8-byte load; the values are the same doubles.
struct GlobalPts {
    const double* p;
    const F4* p32;
    bool use32;
    __device__ P3 operator()(int32_t i) const {
        if (use32) {
            const F4 v = p32[i];
            return P3{(double)v.x, (double)v.y, (double)v.z};
        }
        // (a flat load: measured faster here than the global form, 0.49 vs
        // 0.55 ms per 2^24 in_sphere over an L2-resident set)
        return load_p3(p, i);
    }
};
struct LdsPts {
    const double* p;  // points to the __shared__ array
    // i < kLdsPts < 2^24: the byte offset from the full-rate 24-bit multiply
    __device__ P3 operator()(int32_t i) const {
        const double* q = reinterpret_cast<const double*>(reinterpret_cast<const char*>(p) +
                                                          __umul24((uint32_t)i, 24u));
        return P3{q[0], q[1], q[2]};
    }
};

// The fast tiers of one query: T0 filter, T1 certificate, in-register
// perturbation. Returns the public sign, or 2 when the exact tiers must decide.
// kReload: fetch the points again for T1 instead of keeping them live across
// the filter (cheap from LDS; it lowers register pressure on the common path).
// cert_first (uniform per task, from set_certifiable): the set-level T1
// certificate on the bare determinant first; a certified determinant is exact,
// so its sign is the one the filter or the exact tiers would return.
template <bool kReload, class Fetch>
__device__ __forceinline__ int fast_decide(const Fetch& F, uint8_t o, const int32_t id[5], SetGrid sg,
                                           bool cert_first, int& tier, bool& perturbed) {
    const bool ins = o != DEL3D_OP_ORIENT3D;
    P3 p[5];
    p[0] = F(id[0]);
    p[1] = F(id[1]);
    p[2] = F(id[2]);
    p[3] = F(id[3]);
    if (cert_first) {
        if (!ins) {
            const double md = max_diff(p, 4);
            if (t1_certify_set(md, 15, -358, sg)) {
                tier = kTierCert;
                return -sgn(orient_det(p[0], p[1], p[2], p[3]));
            }
        } else {
            p[4] = F(id[4]);
            const double md = max_diff(p, 5);
            if (t1_certify_set(md, 8, -214, sg)) {
                tier = kTierCert;
                int sign = -sgn(insphere_det(p[0], p[1], p[2], p[3], p[4]));
                if (sign == 0 && o == DEL3D_OP_IN_SPHERE) {
                    perturbed = true;
                    sign = perturbation_walk(p, id, [&](int k) {
                        P3 t[4];
                        subst_tet(p, k, t);
                        return -sgn(orient_det(t[0], t[1], t[2], t[3]));
                    });
                }
                return sign;
            }
        }
    }
    double det;
    if (!ins) {
        const Filt f = orient_filter(p[0], p[1], p[2], p[3]);
        if (filter_decides(f, kOrientBound)) {
            tier = kTierFilter;
            return -sgn(f.det);
        }
        det = f.det;
    } else {
        p[4] = F(id[4]);
        const Filt f = insphere_filter(p[0], p[1], p[2], p[3], p[4]);
        if (filter_decides(f, kInSphereBound)) {
            tier = kTierFilter;
            return -sgn(f.det);
        }
        det = f.det;
    }
    if (kReload) {
        asm volatile("" ::: "memory");  // the points are fetched again, not kept
#pragma unroll
        for (int k = 0; k < 4; ++k) p[k] = F(id[k]);
        if (ins) p[4] = F(id[4]);
    }
    // T1: the set-level grid summary first, the per-coordinate test if it fails
    if (!ins) {
        const double md = max_diff(p, 4);
        if (!t1_certify_set(md, 15, -358, sg) && !t1_certify(p, 4, md, 15, -358)) return 2;
        tier = kTierCert;
        return -sgn(det);  // certified: the filter's determinant is exact
    }
    const double md = max_diff(p, 5);
    if (!t1_certify_set(md, 8, -214, sg) && !t1_certify(p, 5, md, 8, -214)) return 2;
    tier = kTierCert;
    int sign = -sgn(det);
    if (sign == 0 && o == DEL3D_OP_IN_SPHERE) {
        perturbed = true;
        // every four of these points certify at degree 3 too, so the FP64
        // orient3d determinant of a substituted tetrahedron is exact here
        sign = perturbation_walk(p, id, [&](int k) {
            P3 t[4];
            subst_tet(p, k, t);
            return -sgn(orient_filter(t[0], t[1], t[2], t[3]).det);
        });
    }
    return sign;
}

// The points of set s.
__device__ inline const double* set_points(const BatchArgs& a, int32_t s) {
    if (a.set_ptr) return a.set_ptr[s];
    return a.xyz + 3 * (a.pt_off ? a.pt_off[s] : 0);
}

// Task `t` of a batch: explicit table, or implicit equal chunks of one set.
__device__ inline PredTask task_at(const BatchArgs& a, uint32_t t) {
    if (a.tasks) return a.tasks[t];
    PredTask r;
    r.q0 = (int64_t)t * a.chunk;
    r.q1 = min(r.q0 + (int64_t)a.chunk, a.n_q);
    r.set = 0;
    r.pad = 0;
    return r;
}

// Where a fast kernel's points come from. kGlobal: gathers from global
// memory; kLdsF64: the workgroup loads each task's point set into LDS first;
// kLdsPacked: the same with the set's packed copy; kGlobalPacked: gathers of
// the packed copy's words from global memory.
enum { kGlobal = 0, kLdsF64 = 1, kLdsPacked = 2, kGlobalPacked = 3 };

// The per-task constants every query of a task shares.
struct TaskView {
    uint32_t np32;        // indices below this are inside the set
    SetGrid sg;           // the set's grid summary
    bool cert_first;      // set-level T1 certificate first (set_certifiable)
    const double* setpts; // the set's points in global memory
    const F4* set_pts32;  // their float copy, or nullptr
    bool use32;
    const double* s_pts;  // the LDS copy (kLdsF64)
    const uint32_t* s_pk; // the LDS packed copy (kLdsPacked)
};

// One query: index and op checks, the fast tiers, then the sign or a queue
// entry for the exact tiers. Every lane of the wavefront calls it together
// (act: the lane holds a query).
template <bool kStats, int kMode, class Queue>
__device__ __forceinline__ void decide_query(const BatchArgs& a, const TaskView& tv, bool act, int64_t q,
                                             uint8_t o, const int32_t id[5], Queue& wq, WaveStats& ws) {
    const int nchk = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
    bool valid = o <= DEL3D_OP_IN_SPHERE_UNPERTURBED;
#pragma unroll
    for (int k = 0; k < 5; ++k) valid = valid && (k >= nchk || (uint32_t)id[k] < tv.np32);

    int sign = 0, tier = kTierFilter;
    bool perturbed = false;
    if (act && valid) {
        if (kMode == kLdsF64)
            sign = fast_decide<true>(LdsPts{tv.s_pts}, o, id, tv.sg, tv.cert_first, tier, perturbed);
        else if (kMode == kLdsPacked) {
            uint32_t w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) w[k] = k < nchk ? tv.s_pk[id[k]] : 0u;
            sign = packed_decide(w, o, id, tier, perturbed);
        } else if (kMode == kGlobalPacked) {
            uint32_t w[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) w[k] = k < nchk ? a.pk[id[k]] : 0u;
            sign = packed_decide(w, o, id, tier, perturbed);
        } else if (tv.cert_first) {
            sign = fast_decide<false>(GlobalPts{tv.setpts, tv.set_pts32, tv.use32}, o, id, tv.sg,
                                      tv.cert_first, tier, perturbed);
        } else {
            // a set the grid summary cannot certify (arbitrary doubles): the
            // filter rarely fails, so T1 gathers the points again rather than
            // keeping them live across the filter, which would spill a
            // coordinate to scratch on every query (the branch is uniform)
            sign = fast_decide<true>(GlobalPts{tv.setpts, tv.set_pts32, tv.use32}, o, id, tv.sg,
                                     false, tier, perturbed);
        }
    }
    const bool queued = act && valid && sign == 2;
    if (act && !queued) a.out[q] = (int8_t)sign;
    wq.push(queued, (uint32_t)q, &a.ctr->queued, a.queue);
    if (kStats) {
        ws.count(0, act && valid && !queued && tier == kTierFilter);
        ws.count(1, act && valid && !queued && tier == kTierCert);
        ws.count(2, act && perturbed);
        ws.count(3, act && !valid);
    }
}

// Takes the next task for thread 0 and publishes it in s_task: {q0, q1, set,
// points of the set}, q0 = -1 when none is left.
__device__ inline void take_task(const BatchArgs& a, uint32_t* next, int64_t* s_task, uint32_t home,
                                 uint32_t& bin_k, bool binned) {
    uint32_t t = 0xFFFFFFFFu;
    if (binned) {
        // each bin's cursor only grows, so every task of a bin is handed out
        // once; a bin found empty is not tried again
        while (bin_k < kBins) {
            const uint32_t b = (home + bin_k) % kBins;
            const uint32_t lo = a.bins[b], hi = a.bins[b + 1];
            const uint32_t i = atomicAdd(&a.ctr->bin_next[b], 1u);
            if (lo <= hi && i < hi - lo) {
                t = lo + i;
                break;
            }
            ++bin_k;
        }
    } else {
        t = atomicAdd(next, 1u);
    }
    if (t < a.n_tasks) {
        const PredTask tk = task_at(a, t);
        int64_t np = a.n_pts;
        if (a.set_ptr) np = a.set_np[tk.set];
        else if (a.pt_off) np = a.pt_off[tk.set + 1] - a.pt_off[tk.set];
        s_task[0] = tk.q0;
        s_task[1] = tk.q1;
        s_task[2] = tk.set;
        s_task[3] = np;
    } else {
        s_task[0] = -1;
    }
}

// The per-task constants of task (set, np); np already capped to what the
// kernel's copy of the set holds.
template <int kMode>
__device__ __forceinline__ TaskView task_view(const BatchArgs& a, int32_t set, int64_t np, const double* s_pts,
                                              const uint32_t* s_pk) {
    TaskView tv;
    // the same bound for 32-bit compares: an index as uint32 is below np
    // exactly when it is below min(np, 2^31) (negative ones never are)
    tv.np32 = (uint32_t)min(np, (int64_t)0x80000000);
    tv.setpts = set_points(a, set);
    // no summary: an impossible grid, so only the per-coordinate test runs
    tv.sg = SetGrid{-100000, 100000};
    if (a.grid) {
        tv.sg.gmin = __builtin_amdgcn_readfirstlane(a.grid[2 * set]);
        tv.sg.mexp = __builtin_amdgcn_readfirstlane(a.grid[2 * set + 1]);
    }
    tv.cert_first = set_certifiable(tv.sg);
    tv.use32 = kMode == kGlobal && a.pts32 != nullptr && set_f32_exact(tv.sg);
    tv.set_pts32 = tv.use32 ? a.pts32 + (a.pt_off ? a.pt_off[set] : 0) : nullptr;
    tv.s_pts = s_pts;
    tv.s_pk = s_pk;
    return tv;
}

// ---------------------------------------------------------------------------
// Global-gather kernels: 256-thread workgroups, several per CU, 256-query
// tiles staged in LDS and regrouped by op type across the workgroup.
// ---------------------------------------------------------------------------
constexpr int kGlobalQBuf = 1024;  // exact-queue buffer entries per wavefront
struct GlobalLds {
    uint32_t qbuf[kGlobalThreads / 64][kGlobalQBuf];
    int32_t idx[kGlobalThreads * 5];
    uint8_t op[kGlobalThreads];
    uint16_t perm[kGlobalThreads];
    uint32_t wave_ins;  // in_sphere count of each wavefront, one byte each
    int64_t task[4];
};

template <bool kStats, int kMode>
__device__ __forceinline__ void pred_global_body(const BatchArgs& a, GlobalLds& L) {
    constexpr bool kPacked = kMode == kGlobalPacked;
    constexpr int kT = kGlobalThreads;
    constexpr int kW = kT / 64;
    int32_t* s_idx = L.idx;
    uint8_t* s_op = L.op;
    uint16_t* s_perm = L.perm;
    uint8_t* s_wave_ins = reinterpret_cast<uint8_t*>(&L.wave_ins);
    int64_t* s_task = L.task;

    // a single set run by both a packed kernel and the global one: the set's
    // summary picks one, the other returns at once (a.grid holds the summary)
    if (a.pack_route) {
        const bool packable = set_packable(SetGrid{a.grid[0], a.grid[1]});
        if (packable != kPacked) return;
    }

    const int tid = threadIdx.x;
    const int wave = (int)__builtin_amdgcn_readfirstlane((uint32_t)tid >> 6);
    const int lane = __lane_id();
    // thread 0's position in the bin order (own bin first), when binned
    const uint32_t home = blockIdx.x % kBins;
    uint32_t bin_k = 0;
    WaveQueue<kGlobalQBuf> wq{L.qbuf[wave], 0};
    WaveStats ws;

    for (;;) {
        __syncthreads();  // the previous task is done with every LDS array
        if (tid == 0) take_task(a, &a.ctr->next_task[0], s_task, home, bin_k, a.bins != nullptr);
        __syncthreads();
        // the task is the same for every lane: keep it in scalar registers
        const int64_t q0 = uniform64(s_task[0]);
        if (q0 < 0) break;
        const int64_t q1 = uniform64(s_task[1]);
        const int32_t set = (int32_t)uniform64(s_task[2]);
        const TaskView tv = task_view<kMode>(a, set, uniform64(s_task[3]), nullptr, nullptr);

        // Each thread's share of a tile's records (dwords tid + k*kT of the
        // 5*kT, and op byte tid) is loaded one tile ahead into registers, so
        // the load latency overlaps the current tile's work.
        int32_t pf_idx[5];
        uint8_t pf_op = (uint8_t)DEL3D_OP_ORIENT3D;
        auto prefetch = [&](int64_t b) {
            const int c = (int)min((int64_t)kT, q1 - b);
            const int32_t* src = a.idx + b * 5;
            if (c == kT) {  // a full tile (uniform): no per-lane guards
#pragma unroll
                for (int k = 0; k < 5; ++k) pf_idx[k] = src[tid + k * kT];
                pf_op = a.op[b + tid];
                return;
            }
#pragma unroll
            for (int k = 0; k < 5; ++k) {
                const int j = tid + k * kT;
                pf_idx[k] = j < c * 5 ? src[j] : 0;
            }
            pf_op = tid < c ? a.op[b + tid] : (uint8_t)DEL3D_OP_ORIENT3D;
        };
        prefetch(q0);
        // Two barriers per tile. The counts' byte is written before the first;
        // its readers all passed the previous tile's second barrier. The
        // records, s_perm and s_op are written between the two: every thread
        // that reaches the first has finished the previous tile.
        for (int64_t base = q0; base < q1; base += kT) {
            const uint32_t cnt = (uint32_t)min((int64_t)kT, q1 - base);
            const bool live = (uint32_t)tid < cnt;
            const uint8_t my_op = live ? pf_op : (uint8_t)DEL3D_OP_ORIENT3D;

            // regroup: in_sphere-type queries first, orient3d after, each in
            // query order, so wavefronts mostly run one code path
            const bool is_ins = live && my_op != DEL3D_OP_ORIENT3D;
            const uint64_t ins_mask = __ballot(is_ins);
            if (lane == 0) s_wave_ins[wave] = (uint8_t)__popcll(ins_mask);  // <= 64
            __syncthreads();
#pragma unroll
            for (int k = 0; k < 5; ++k) s_idx[tid + k * kT] = pf_idx[k];
            if (base + kT < q1) prefetch(base + kT);
            // sums of the byte counts: all of them, and those of the
            // wavefronts before this one
            const uint32_t d = L.wave_ins;
            const uint32_t m = wave >= 4 ? 0x01010101u : wave <= 0 ? 0u : (0x01010101u >> (8 * (4 - wave)));
            const uint32_t ins_total = __builtin_amdgcn_udot4(d, 0x01010101u, 0u, false);
            const uint32_t ins_before = __builtin_amdgcn_udot4(d, m, 0u, false);
            static_assert(kW == 4, "one dword of byte counts");
            const uint32_t lt = (uint32_t)__popcll(ins_mask & lanemask_lt());
            const uint32_t pos = is_ins ? ins_before + lt
                                        : ins_total + ((uint32_t)tid - (ins_before + lt));
            if (live) s_perm[pos] = (uint16_t)tid;
            s_op[tid] = my_op;
            __syncthreads();

            const bool act = (uint32_t)tid < cnt;
            const int lq = act ? s_perm[tid] : 0;
            int32_t id[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) id[k] = s_idx[lq * 5 + k];
            decide_query<kStats, kMode>(a, tv, act, base + lq, s_op[lq], id, wq, ws);
        }
    }
    wq.flush(&a.ctr->queued, a.queue);
    if (kStats) ws.flush(a.stats);
}

// ---------------------------------------------------------------------------
// LDS kernels: one 1024-thread workgroup per CU holds a task's point set in
// LDS; its 16 wavefronts then run independently, with no barrier until the
// task ends. A wavefront takes chunks of 64*kq consecutive queries (kq of
// them per lane, 1 <= kq <= 4) from a counter in LDS, regroups each chunk so
// that in_sphere-type queries fill whole 64-query packs first and orient3d
// queries the packs after (at most one pack of a chunk mixes both), and runs
// the packs. The regrouping goes through a per-wavefront table in LDS: the
// position in the regrouped order of each query is its rank among those of
// its kind, and the table maps positions back to queries (9 bits) and their
// op (2 bits; 3 = an unknown op byte). Loads run ahead of the arithmetic: a
// pack's index records are loaded while the pack before it is decided, and
// a chunk's op bytes while the chunk before it runs.
// ---------------------------------------------------------------------------
constexpr int kLdsWaves = kLdsThreads / 64;
constexpr int kChunkMax = 256;
constexpr int kChunkCols = kChunkMax / 64;
constexpr int kLdsQBuf = 256;  // exact-queue buffer entries per wavefront

template <int kMode>
struct WaveLds {
    double pts[kMode == kLdsF64 ? 3 * kLdsPts : 1];
    uint32_t pk[kMode == kLdsPacked ? kPackPts : 1];
    uint32_t qbuf[kLdsWaves][kLdsQBuf];
    uint16_t sel[kLdsWaves][kChunkMax];
    int64_t task[4];
    uint32_t next_chunk;
};

template <bool kStats, int kMode>
__device__ __forceinline__ void pred_lds_body(const BatchArgs& a, WaveLds<kMode>& L) {
    constexpr int kCap = kMode == kLdsF64 ? kLdsPts : kPackPts;
    if (kMode == kLdsPacked && a.pack_route) {
        if (!set_packable(SetGrid{a.grid[0], a.grid[1]})) return;
    }
    const int tid = threadIdx.x;
    const int wave = (int)__builtin_amdgcn_readfirstlane((uint32_t)tid >> 6);
    const int lane = __lane_id();
    uint16_t* sel = L.sel[wave];
    int32_t loaded_set = -1;
    uint32_t bin_k = 0;
    WaveQueue<kLdsQBuf> wq{L.qbuf[wave], 0};
    WaveStats ws;

    for (;;) {
        __syncthreads();  // every wavefront is done with the previous task
        if (tid == 0) {
            take_task(a, &a.ctr->next_task[1], L.task, 0, bin_k, false);
            L.next_chunk = 0;
        }
        __syncthreads();
        const int64_t q0 = uniform64(L.task[0]);
        if (q0 < 0) break;
        const int64_t q1 = uniform64(L.task[1]);
        const int32_t set = (int32_t)uniform64(L.task[2]);
        // the host routes only sets of at most kCap points here
        const int64_t np = min(uniform64(L.task[3]), (int64_t)kCap);
        if (set != loaded_set) {
            if (kMode == kLdsF64) {
                const double* src = set_points(a, set);
                const int n3 = 3 * (int)np;
                for (int i = tid; i < n3; i += kLdsThreads) L.pts[i] = src[i];
            } else {
                for (int i = tid; i < (int)np; i += kLdsThreads) L.pk[i] = a.pk[i];
            }
            loaded_set = set;
            __syncthreads();
        }
        const TaskView tv = task_view<kMode>(a, set, np, L.pts, L.pk);

        // queries per lane of a chunk: the most that still gives every
        // wavefront two chunks of the task
        const int64_t len = q1 - q0;
        int kq = kChunkCols;
        while (kq > 1 && len < (int64_t)kLdsWaves * 64 * kq * 2) kq >>= 1;
        const int csz = 64 * kq;

        // the first query of this wavefront's next chunk, q1 when none is left
        auto claim = [&]() -> int64_t {
            uint32_t c = 0;
            if (lane == 0) c = atomicAdd(&L.next_chunk, 1u);
            c = __builtin_amdgcn_readfirstlane(__shfl(c, 0));
            return min(q0 + (int64_t)c * csz, q1);
        };
        // The loads below run ahead of the arithmetic and are issued
        // unconditionally (inactive lanes read a query of the task), so that
        // every pack issues the same loads in the same order and a pack's
        // wait for its records does not also wait for the loads after them.
        // the op bytes of the chunk at b: query 64k + lane in ops[k]
        // (a uniform base and 32-bit lane offsets: no 64-bit address
        // arithmetic per lane; past the task's end every lane reads its last
        // byte)
        uint8_t ops[kChunkCols];
        auto load_ops = [&](int64_t b) {
            const int64_t bb = min(b, q1 - 1);
            const uint8_t* ob = a.op + bb;
            const uint32_t lim = (uint32_t)min(q1 - 1 - bb, (int64_t)(kChunkMax - 1));
#pragma unroll
            for (int k = 0; k < kChunkCols; ++k) ops[k] = ob[min((uint32_t)(64 * k + lane), lim)];
        };
        // the table of the chunk at b, from ops; returns its query count
        auto regroup = [&](int64_t b) -> int {
            const int cnt = (int)min((int64_t)csz, q1 - b);
            uint64_t insm[kChunkCols];
            int n_ins = 0;
#pragma unroll
            for (int k = 0; k < kChunkCols; ++k) {
                insm[k] = 0;
                if (k < kq) {
                    insm[k] = __ballot(64 * k + lane < cnt && ops[k] != DEL3D_OP_ORIENT3D);
                    n_ins += __popcll(insm[k]);
                }
            }
            int before = 0;  // in_sphere-type queries of the columns before k
#pragma unroll
            for (int k = 0; k < kChunkCols; ++k) {
                if (k < kq) {
                    const int j = 64 * k + lane;
                    const int lt = before + __popcll(insm[k] & lanemask_lt());
                    const bool is_ins = (insm[k] >> lane) & 1ull;
                    const int pos = is_ins ? lt : n_ins + (j - lt);
                    const uint32_t ob = ops[k] > DEL3D_OP_IN_SPHERE_UNPERTURBED ? 3u : (uint32_t)ops[k];
                    if (j < cnt) sel[pos] = (uint16_t)((uint32_t)j | (ob << 9));
                    before += __popcll(insm[k]);
                }
            }
            wave_lds_order();
            return cnt;
        };
        // pack k of the chunk at b (cnt queries): the lane's table entry,
        // and the loads of its query's index record
        auto read_pack = [&](int64_t b, int k, int cnt, uint32_t& s, int32_t (&id)[5]) {
            const bool act = 64 * k + lane < cnt;
            s = act ? sel[64 * k + lane] : 0u;
            const int32_t* rec = a.idx + b * 5 + __umul24(s & 511u, 5u);
#pragma unroll
            for (int m = 0; m < 5; ++m) id[m] = rec[m];
        };

        int64_t cb = claim();  // the chunk of the pack being decided
        if (cb >= q1) continue;
        load_ops(cb);
        int ccnt = regroup(cb);
        int64_t nb = claim();  // the chunk after it
        load_ops(nb);
        int k = 0;
        uint32_t s;
        int32_t id[5];
        read_pack(cb, 0, ccnt, s, id);
        for (;;) {
            // the pack after this one: its loads go out before this one's work
            int64_t xb = cb;
            int xk = k + 1, xcnt = ccnt;
            bool more = true;
            if (64 * xk >= ccnt) {
                if (nb < q1) {
                    // this chunk's table is read: the next chunk's replaces it
                    xb = nb;
                    xk = 0;
                    xcnt = regroup(nb);
                    nb = claim();
                } else {
                    more = false;
                    xk = k;  // (loads of this pack again, unused)
                }
            }
            uint32_t ns;
            int32_t nid[5];
            read_pack(xb, xk, xcnt, ns, nid);
            load_ops(nb);

            const bool act = 64 * k + lane < ccnt;
            const uint32_t ob = s >> 9;
            decide_query<kStats, kMode>(a, tv, act, cb + (int64_t)(s & 511u),
                                        ob == 3u ? (uint8_t)0xFF : (uint8_t)ob, id, wq, ws);
            if (!more) break;
            cb = xb;
            k = xk;
            ccnt = xcnt;
            s = ns;
#pragma unroll
            for (int m = 0; m < 5; ++m) id[m] = nid[m];
        }
    }
    wq.flush(&a.ctr->queued, a.queue);
    if (kStats) ws.flush(a.stats);
}

template <bool kStats>
__global__ __launch_bounds__(kLdsThreads) void pred_batch_lds(BatchArgs a) {
    __shared__ WaveLds<kLdsF64> L;
    pred_lds_body<kStats, kLdsF64>(a, L);
}

template <bool kStats>
__global__ __launch_bounds__(kLdsThreads) void pred_batch_packed(BatchArgs a) {
    __shared__ WaveLds<kLdsPacked> L;
    pred_lds_body<kStats, kLdsPacked>(a, L);
}

// The global kernels are held to 128 VGPRs (four wavefronts per SIMD) to
// hide gather latency; the LDS ones get that bound from their 1024-thread
// workgroup.
template <bool kStats>
__global__ __launch_bounds__(kGlobalThreads) __attribute__((amdgpu_waves_per_eu(4)))
void pred_batch_global(BatchArgs a) {
    __shared__ GlobalLds L;
    pred_global_body<kStats, kGlobal>(a, L);
}

template <bool kStats>
__global__ __launch_bounds__(kGlobalThreads) __attribute__((amdgpu_waves_per_eu(4)))
void pred_batch_global_packed(BatchArgs a) {
    __shared__ GlobalLds L;
    pred_global_body<kStats, kGlobalPacked>(a, L);
}

// ---------------------------------------------------------------------------
// Exact path, registers (1..4 limbs)
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

// The points of the set holding query q: binary search over the query
// offsets (one set: set 0).
__device__ inline const double* points_of_query(const BatchArgs& a, int64_t q) {
    if (!a.q_off) return set_points(a, 0);
    uint32_t lo = 0, hi = a.n_sets;  // last s with q_off[s] <= q
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.q_off[mid] <= q) lo = mid;
        else hi = mid;
    }
    return set_points(a, (int32_t)lo);
}

__global__ __launch_bounds__(256) void pred_exact(BatchArgs a) {
    const uint32_t n = a.ctr->queued;
    const uint32_t stride = gridDim.x * blockDim.x;
    // i0 is the wavefront's base, so the trip count is uniform per wavefront
    // and the ballots in wave_count see every lane
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        const bool act = i < n;
        bool perturbed = false, done = false;
        if (act) {
            const uint32_t q = a.queue[i];
            const uint8_t o = a.op[q];
            const double* pts = points_of_query(a, q);
            int32_t id[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) id[k] = a.idx[(size_t)q * 5 + k];
            P3 p[5];
            const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
            // all five slots with static indices (an orient3d query repeats
            // its last point, which leaves the grid unchanged), so p[] stays
            // in registers
#pragma unroll
            for (int k = 0; k < 5; ++k) p[k] = load_p3(pts, id[k < np ? k : 3]);
            const Grid g = grid_of(p, 5);
            const int L = limbs_for(g);
            int s = 0;
            done = true;
            if (L == 1) {
                s = exact_query<1>(o, p, id, g.q, perturbed);
            } else if (L == 2) {
                s = exact_query<2>(o, p, id, g.q, perturbed);
            } else {
                done = false;
                a.wide[atomicAdd(&a.ctr->wide, 1u)] = q;  // rare: exponent span above 2^55
            }
            if (done) a.out[q] = (int8_t)s;
        }
        if (a.stats) {
            wave_count(act && done, &a.stats[2]);
            wave_count(act && perturbed, &a.stats[4]);
        }
    }
}

// The wide queue's queries of 3 or 4 limbs per difference (exponent spans up
// to 2^119: random doubles whose coordinates differ much in magnitude), in
// registers like pred_exact; its own kernel, so that pred_exact keeps the
// occupancy of the 1-2 limb forms. Wider queries are left to pred_wide.
__global__ __launch_bounds__(256) void pred_exact_mid(BatchArgs a) {
    const uint32_t n = a.ctr->wide;
    const uint32_t stride = gridDim.x * blockDim.x;
    for (uint32_t i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63u); i0 < n; i0 += stride) {
        const uint32_t i = i0 + __lane_id();
        bool perturbed = false, done = false;
        if (i < n) {
            const uint32_t q = a.wide[i];
            const uint8_t o = a.op[q];
            const double* pts = points_of_query(a, q);
            int32_t id[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) id[k] = a.idx[(size_t)q * 5 + k];
            P3 p[5];
            const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
            // all five slots with static indices (an orient3d query repeats
            // its last point, which leaves the grid unchanged), so p[] stays
            // in registers
#pragma unroll
            for (int k = 0; k < 5; ++k) p[k] = load_p3(pts, id[k < np ? k : 3]);
            const Grid g = grid_of(p, 5);
            const int L = limbs_for(g);
            int s = 0;
            if (L == 3) {
                s = exact_query<3>(o, p, id, g.q, perturbed);
                done = true;
            } else if (L == 4) {
                s = exact_query<4>(o, p, id, g.q, perturbed);
                done = true;
            }
            if (done) a.out[q] = (int8_t)s;
        }
        if (a.stats) {
            wave_count(done, &a.stats[2]);
            wave_count(perturbed, &a.stats[4]);
        }
    }
}

// ---------------------------------------------------------------------------
// Small batches of 3D queries in one pass
// ---------------------------------------------------------------------------
// One lane per query runs every tier the batch pipeline runs (the filter, the
// per-coordinate certificate, exact integers of 1-4 limbs in registers, the
// perturbation); the rare wider query goes to the wide queue for pred_wide.
// No grid summary pass, no exact queue: for batches of a few thousand
// queries, where the pipeline's launches cost more than its work (a driver's
// conflict-test waves).
__global__ __launch_bounds__(256) void pred_small(BatchArgs a) {
    const uint32_t np32 = (uint32_t)min(a.n_pts, (int64_t)0x80000000);
    const SetGrid none{-100000, 100000};  // no set summary: the per-coordinate certificate only
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < a.n_q;
         q += (int64_t)gridDim.x * blockDim.x) {
        const uint8_t o = a.op[q];
        int32_t id[5];
#pragma unroll
        for (int k = 0; k < 5; ++k) id[k] = a.idx[5 * q + k];
        const int nchk = o == DEL3D_OP_ORIENT3D ? 4 : 5;
        bool valid = o <= DEL3D_OP_IN_SPHERE_UNPERTURBED;
#pragma unroll
        for (int k = 0; k < 5; ++k) valid = valid && (k >= nchk || (uint32_t)id[k] < np32);
        if (!valid) {
            a.out[q] = 0;
            continue;
        }
        int tier = kTierFilter;
        bool perturbed = false;
        int sign = fast_decide<false>(GlobalPts{a.xyz, nullptr, false}, o, id, none, false, tier, perturbed);
        if (sign == 2) {
            P3 p[5];
#pragma unroll
            for (int k = 0; k < 5; ++k) p[k] = load_p3(a.xyz, id[k < nchk ? k : 3]);
            const Grid g = grid_of(p, 5);  // the repeated point leaves the grid unchanged
            const int L = limbs_for(g);
            if (L == 1) sign = exact_query<1>(o, p, id, g.q, perturbed);
            else if (L == 2) sign = exact_query<2>(o, p, id, g.q, perturbed);
            else if (L == 3) sign = exact_query<3>(o, p, id, g.q, perturbed);
            else if (L == 4) sign = exact_query<4>(o, p, id, g.q, perturbed);
            else {
                a.wide[atomicAdd(&a.ctr->wide, 1u)] = (uint32_t)q;  // pred_wide writes it
                continue;
            }
        }
        a.out[q] = (int8_t)sign;
    }
}

// ---------------------------------------------------------------------------
// Exact path, any width, LDS: one query per workgroup, its products spread
// over the lanes
// ---------------------------------------------------------------------------
// The determinants' integer products form a short DAG (differences -> 2x2
// minors -> 3x3 cofactors and lifts -> the 4x4 expansion), and the products of
// one level are independent: each runs serially on its own lane, in its own
// LDS slots, with a barrier between levels. The critical path is then one
// product per level (about 9 L^2 limb multiplies for in_sphere, against 72 L^2
// for the whole determinant on one lane). Every lane of the workgroup runs the
// same control flow (the query, its limb count and the perturbation walk are
// uniform), so the barriers inside are reached by all of them.

// LDS limbs of one workgroup's workspace, per limb of L.
constexpr int kCoopPerLimb = 3 + 12 + 12 + 12 + 8 + 24 * 3 + 24 * 3;  // 191

struct CoopWs {
    uint32_t* base;
    int L;
    __device__ uint32_t* ev() const { return base; }                       // 3L: the base point
    __device__ uint32_t* diff(int i, int k) const { return base + (3 + 3 * i + k) * L; }  // 12L
    __device__ uint32_t* minor(int k) const { return base + (15 + 2 * k) * L; }          // 6 x 2L
    __device__ uint32_t* cof(int k) const { return base + (27 + 3 * k) * L; }            // 4 x 3L
    __device__ uint32_t* lift(int k) const { return base + (39 + 2 * k) * L; }           // 4 x 2L
    __device__ uint32_t* slot(int k) const { return base + (47 + 3 * k) * L; }           // 24 x 3L
    __device__ uint32_t* tmp(int k) const { return base + (47 + 72 + 3 * k) * L; }       // 24 x 3L
    // the four 5L products of the last level, and their temporaries, reuse
    // the slot and temporary areas: 4 x 5L <= 72L each
    __device__ uint32_t* wide_slot(int k) const { return base + (47 + 5 * k) * L; }
    __device__ uint32_t* wide_tmp(int k) const { return base + (47 + 72 + 5 * k) * L; }
};

// Sign of |a-d, b-d, c-d| (the computed orient3d form), cooperatively.
__device__ int coop_orient_raw(const P3& a, const P3& b, const P3& c, const P3& d, int q, const CoopWs& W,
                               int* s_sign) {
    const int t = threadIdx.x, L = W.L;
    const P3* pt[3] = {&a, &b, &c};
    if (t < 3) rt_to_limbs(t == 0 ? d.x : t == 1 ? d.y : d.z, q, W.ev() + t * L, L);
    __syncthreads();
    if (t < 9) {
        const int i = t / 3, k = t % 3;
        uint32_t* r = W.diff(i, k);
        rt_to_limbs(k == 0 ? pt[i]->x : k == 1 ? pt[i]->y : pt[i]->z, q, r, L);
        rt_addsub(r, r, W.ev() + k * L, L, 1);
    }
    __syncthreads();
    // minors m_k = X(i) Y(j) - X(j) Y(i) for (i, j) = (b, c), (c, a), (a, b)
    const int mi[3][2] = {{1, 2}, {2, 0}, {0, 1}};
    if (t < 6) {
        const int k = t >> 1, second = t & 1;
        const int i = mi[k][second], j = mi[k][1 - second];
        rt_mul(W.slot(t), W.diff(i, 0), L, W.diff(j, 1), L, W.tmp(t));
    }
    __syncthreads();
    if (t < 3) rt_addsub(W.minor(t), W.slot(2 * t), W.slot(2 * t + 1), 2 * L, 1);
    __syncthreads();
    if (t < 3) rt_mul(W.slot(t), W.diff(t, 2), L, W.minor(t), 2 * L, W.tmp(t));
    __syncthreads();
    if (t == 0) {
        uint32_t* det = W.slot(0);
        rt_addsub(det, det, W.slot(1), 3 * L, 0);
        rt_addsub(det, det, W.slot(2), 3 * L, 0);
        *s_sign = rt_sign(det, 3 * L);
    }
    __syncthreads();
    const int sign = *s_sign;
    __syncthreads();  // s_sign is reused by the next call
    return sign;
}

// Sign of the lifted determinant relative to e (the computed in_sphere form),
// cooperatively; the same products as rt_insphere_raw (src/predicates.cpp:
// 165-176's cofactors and lifts).
__device__ int coop_insphere_raw(const P3& a, const P3& b, const P3& c, const P3& d, const P3& e, int q,
                                 const CoopWs& W, int* s_sign) {
    const int t = threadIdx.x, L = W.L;
    const P3* pt[4] = {&a, &b, &c, &d};
    if (t < 3) rt_to_limbs(t == 0 ? e.x : t == 1 ? e.y : e.z, q, W.ev() + t * L, L);
    __syncthreads();
    if (t < 12) {
        const int i = t / 3, k = t % 3;
        uint32_t* r = W.diff(i, k);
        rt_to_limbs(k == 0 ? pt[i]->x : k == 1 ? pt[i]->y : pt[i]->z, q, r, L);
        rt_addsub(r, r, W.ev() + k * L, L, 1);
    }
    __syncthreads();
    // minors ab bc cd da ac bd: X(i) Y(j) - X(j) Y(i)
    const int mp[6][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {0, 2}, {1, 3}};
    if (t < 12) {
        const int k = t >> 1, second = t & 1;
        const int i = mp[k][second], j = mp[k][1 - second];
        rt_mul(W.slot(t), W.diff(i, 0), L, W.diff(j, 1), L, W.tmp(t));
    }
    __syncthreads();
    if (t < 6) rt_addsub(W.minor(t), W.slot(2 * t), W.slot(2 * t + 1), 2 * L, 1);
    __syncthreads();
    // cofactors abc bcd cda dab: three terms z(point) * minor each; lifts:
    // x^2 + y^2 + z^2 of each difference
    const int cz_[4][3] = {{0, 1, 2}, {1, 2, 3}, {2, 3, 0}, {3, 0, 1}};
    const int cm_[4][3] = {{1, 4, 0}, {2, 5, 1}, {3, 4, 2}, {0, 5, 3}};
    const int cs_[4][3] = {{0, 1, 0}, {0, 1, 0}, {0, 0, 0}, {0, 0, 0}};  // 1 = subtract
    if (t < 12) {
        const int k = t / 3, j = t % 3;
        rt_mul(W.slot(t), W.diff(cz_[k][j], 2), L, W.minor(cm_[k][j]), 2 * L, W.tmp(t));
    } else if (t < 24) {
        const int k = (t - 12) / 3, j = (t - 12) % 3;
        rt_mul(W.slot(t), W.diff(k, j), L, W.diff(k, j), L, W.tmp(t));
    }
    __syncthreads();
    if (t < 4) {
        uint32_t* ck = W.cof(t);
        for (int i = 0; i < 3 * L; ++i) ck[i] = 0;
        for (int j = 0; j < 3; ++j) rt_addsub(ck, ck, W.slot(3 * t + j), 3 * L, cs_[t][j]);
    } else if (t < 8) {
        const int k = t - 4;
        uint32_t* lk = W.lift(k);
        for (int i = 0; i < 2 * L; ++i) lk[i] = 0;
        for (int j = 0; j < 3; ++j) rt_addsub(lk, lk, W.slot(12 + 3 * k + j), 2 * L, 0);
    }
    __syncthreads();
    // det = (dlift*abc - clift*dab) + (blift*cda - alift*bcd)
    const int dl[4] = {3, 2, 1, 0}, dc[4] = {0, 3, 2, 1};
    if (t < 4) rt_mul(W.wide_slot(t), W.lift(dl[t]), 2 * L, W.cof(dc[t]), 3 * L, W.wide_tmp(t));
    __syncthreads();
    if (t == 0) {
        uint32_t* det = W.wide_slot(0);
        rt_addsub(det, det, W.wide_slot(1), 5 * L, 1);
        rt_addsub(det, det, W.wide_slot(2), 5 * L, 0);
        rt_addsub(det, det, W.wide_slot(3), 5 * L, 1);
        *s_sign = rt_sign(det, 5 * L);
    }
    __syncthreads();
    const int sign = *s_sign;
    __syncthreads();
    return sign;
}

__global__ __launch_bounds__(kWideBlock) void pred_wide(BatchArgs a) {
    __shared__ uint32_t ws[kCoopPerLimb * kWideLimbsMax];
    __shared__ int s_sign;
    const uint32_t n = a.ctr->wide;
    for (uint32_t i = blockIdx.x; i < n; i += gridDim.x) {
        // every lane reads the same query: the control flow below is uniform
        const uint32_t q = a.wide[i];
        const uint8_t o = a.op[q];
        const double* pts = points_of_query(a, q);
        int32_t id[5];
        for (int k = 0; k < 5; ++k) id[k] = a.idx[(size_t)q * 5 + k];
        P3 p[5];
        const int np = (o == DEL3D_OP_ORIENT3D) ? 4 : 5;
        for (int k = 0; k < np; ++k) p[k] = load_p3(pts, id[k]);
        if (np == 4) p[4] = p[3];
        const Grid g = grid_of(p, np);
        const int L = limbs_for(g);
        if (L <= 4) continue;  // decided by pred_exact_mid
        const CoopWs W{ws, L};
        int s;
        if (o == DEL3D_OP_ORIENT3D) {
            s = -coop_orient_raw(p[0], p[1], p[2], p[3], g.q, W, &s_sign);
        } else {
            s = -coop_insphere_raw(p[0], p[1], p[2], p[3], p[4], g.q, W, &s_sign);
            if (s == 0 && o == DEL3D_OP_IN_SPHERE) {
                if (a.stats && threadIdx.x == 0) atomicAdd(&a.stats[4], 1ull);
                s = perturbation_walk(p, id, [&](int k) {
                    P3 t[4];
                    subst_tet(p, k, t);
                    const Filt f = orient_filter(t[0], t[1], t[2], t[3]);
                    if (filter_decides(f, kOrientBound)) return -sgn(f.det);
                    return -coop_orient_raw(t[0], t[1], t[2], t[3], g.q, W, &s_sign);
                });
            }
        }
        if (threadIdx.x == 0) {
            a.out[q] = (int8_t)s;
            if (a.stats) atomicAdd(&a.stats[3], 1ull);
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Set-level grid summaries (SetGrid) for the T1 certificate
// ---------------------------------------------------------------------------
// Also zeroes the batch's counters when given (one launch for both).
__global__ __launch_bounds__(256) void k_grid_init(int32_t* grid, uint32_t n_sets, PredCounters* ctr) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n_sets; i += gridDim.x * blockDim.x) {
        grid[2 * i] = 0x7FFFFFFF;      // min lowest-bit exponent
        grid[2 * i + 1] = -0x7FFFFFFF;  // max magnitude exponent
    }
    if (ctr && blockIdx.x == 0 && threadIdx.x < sizeof(PredCounters) / 4)
        reinterpret_cast<uint32_t*>(ctr)[threadIdx.x] = 0u;
}

// Each workgroup takes one contiguous range of the global point numbering
// (set s owns [first[s], first[s+1])), so a thread's points change set only
// at set boundaries: it accumulates its current set's minimum and maximum in
// registers, flushes them with atomics when the set changes, and at the end
// thread 0 merges the workgroup's runs of equal sets (one atomic pair per
// run). A thread issues kGridUnroll loads before it uses any: doubles of the
// flat coordinate array, or, with out32 (plans: points numbered by a.pt_off),
// whole points, so that the same pass writes the float copy of every point:
// the kernels read it only for a set whose summary says its coordinates are
// exact in float, and there it equals the doubles.
constexpr int kGridUnroll = 4;

__device__ inline void grid_flush(int32_t* grid, int32_t set, int32_t lo, int32_t hi) {
    if (lo <= hi) {
        atomicMin(&grid[2 * set], lo);
        atomicMax(&grid[2 * set + 1], hi);
    }
}

__device__ inline void grid_acc(double c, int32_t& lo, int32_t& hi) {
    const Bits d = decompose(c);
    if (d.m) {
        lo = min(lo, d.e + __builtin_ctzll(d.m));
        hi = max(hi, d.e + 63 - __builtin_clzll(d.m));
    }
}

__global__ __launch_bounds__(256) void k_grid_points(BatchArgs a, uint32_t n_sets, int64_t total,
                                                     const int64_t* __restrict__ first,
                                                     int32_t* __restrict__ grid, F4* __restrict__ out32) {
    __shared__ int32_t s_lo[256], s_hi[256], s_set[256];
    constexpr int64_t kStep = 256 * kGridUnroll;
    int32_t cur = -1, lo = 0x7FFFFFFF, hi = -0x7FFFFFFF;
    int64_t cur_end = 0, cur_first = 0;
    const double* pts = nullptr;
    if (!out32) {
        // summaries only: the coordinates as a flat array of doubles (the
        // summary does not depend on which coordinate a double is), so a
        // wavefront's loads are consecutive doubles; set s owns doubles
        // [3*first[s], 3*first[s+1])
        const int64_t total_d = 3 * total;
        const int64_t per_d = ((total_d + gridDim.x - 1) / gridDim.x + kStep - 1) / kStep * kStep;
        const int64_t d_b0 = (int64_t)blockIdx.x * per_d, d_b1 = min(total_d, d_b0 + per_d);
        for (int64_t d0 = d_b0 + threadIdx.x; d0 < d_b1; d0 += kStep) {
            if (cur >= 0 && d0 + 256 * (kGridUnroll - 1) < min(cur_end, d_b1)) {
                double v[kGridUnroll];
#pragma unroll
                for (int k = 0; k < kGridUnroll; ++k) v[k] = pts[d0 + 256 * k - cur_first];
#pragma unroll
                for (int k = 0; k < kGridUnroll; ++k) grid_acc(v[k], lo, hi);
                continue;
            }
            for (int k = 0; k < kGridUnroll; ++k) {
                const int64_t d = d0 + 256 * k;
                if (d >= d_b1) break;
                if (cur < 0 || d >= cur_end) {
                    grid_flush(grid, cur, lo, hi);
                    lo = 0x7FFFFFFF;
                    hi = -0x7FFFFFFF;
                    if (first) {
                        const int64_t g = d / 3;
                        uint32_t l = cur < 0 ? 0 : (uint32_t)cur, h = n_sets;  // last s with first[s] <= g
                        while (h - l > 1) {
                            const uint32_t mid = (l + h) >> 1;
                            if (first[mid] <= g) l = mid;
                            else h = mid;
                        }
                        cur = (int32_t)l;
                        cur_first = 3 * first[l];
                        cur_end = 3 * first[l + 1];
                    } else {
                        cur = 0;
                        cur_first = 0;
                        cur_end = total_d;
                    }
                    pts = set_points(a, cur);
                }
                grid_acc(pts[d - cur_first], lo, hi);
            }
        }
    }
    const int64_t per = ((total + gridDim.x - 1) / gridDim.x + kStep - 1) / kStep * kStep;
    const int64_t b0 = out32 ? (int64_t)blockIdx.x * per : total, b1 = min(total, b0 + per);
    for (int64_t g0 = b0 + threadIdx.x; g0 < b1; g0 += kStep) {
        if (cur >= 0 && g0 + 256 * (kGridUnroll - 1) < min(cur_end, b1)) {
            // all of this step's points lie in the current set
            double v[kGridUnroll][3];
#pragma unroll
            for (int k = 0; k < kGridUnroll; ++k) {
                const double* p = pts + 3 * (g0 + 256 * k - cur_first);
                v[k][0] = p[0];
                v[k][1] = p[1];
                v[k][2] = p[2];
            }
#pragma unroll
            for (int k = 0; k < kGridUnroll; ++k) {
                grid_acc(v[k][0], lo, hi);
                grid_acc(v[k][1], lo, hi);
                grid_acc(v[k][2], lo, hi);
                if (out32) out32[g0 + 256 * k] = F4{(float)v[k][0], (float)v[k][1], (float)v[k][2], 0.0f};
            }
            continue;
        }
        for (int k = 0; k < kGridUnroll; ++k) {
            const int64_t g = g0 + 256 * k;
            if (g >= b1) break;
            if (cur < 0 || g >= cur_end) {
                grid_flush(grid, cur, lo, hi);
                lo = 0x7FFFFFFF;
                hi = -0x7FFFFFFF;
                if (first) {
                    uint32_t l = cur < 0 ? 0 : (uint32_t)cur, h = n_sets;  // last s with first[s] <= g
                    while (h - l > 1) {
                        const uint32_t mid = (l + h) >> 1;
                        if (first[mid] <= g) l = mid;
                        else h = mid;
                    }
                    cur = (int32_t)l;
                    cur_first = first[l];
                    cur_end = first[l + 1];
                } else {
                    cur = 0;
                    cur_first = 0;
                    cur_end = total;
                }
                pts = set_points(a, cur);
            }
            const double* p = pts + 3 * (g - cur_first);
            const double x = p[0], y = p[1], z = p[2];
            grid_acc(x, lo, hi);
            grid_acc(y, lo, hi);
            grid_acc(z, lo, hi);
            if (out32) out32[g] = F4{(float)x, (float)y, (float)z, 0.0f};
        }
    }
    s_lo[threadIdx.x] = lo;
    s_hi[threadIdx.x] = hi;
    s_set[threadIdx.x] = cur;
    __syncthreads();
    // common case: the whole workgroup saw one set (thread 0 holds the
    // range's first points, so its set is the reference): a tree reduction
    // and one atomic pair
    const int32_t ref = s_set[0];
    if (__syncthreads_and(cur < 0 || cur == ref)) {
        for (int w = blockDim.x / 2; w > 0; w >>= 1) {
            if ((int)threadIdx.x < w) {
                s_lo[threadIdx.x] = min(s_lo[threadIdx.x], s_lo[threadIdx.x + w]);
                s_hi[threadIdx.x] = max(s_hi[threadIdx.x], s_hi[threadIdx.x + w]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0 && ref >= 0) grid_flush(grid, ref, s_lo[0], s_hi[0]);
        return;
    }
    if (threadIdx.x == 0) {
        int32_t c = -1, l = 0x7FFFFFFF, h = -0x7FFFFFFF;
        for (int t = 0; t < (int)blockDim.x; ++t) {
            const int32_t st = s_set[t];
            if (st < 0) continue;
            if (st != c) {
                if (c >= 0) grid_flush(grid, c, l, h);
                c = st;
                l = s_lo[t];
                h = s_hi[t];
            } else {
                l = min(l, s_lo[t]);
                h = max(h, s_hi[t]);
            }
        }
        if (c >= 0) grid_flush(grid, c, l, h);
    }
}

// The float copy of the points of every set whose summary (computed just
// before, on the same stream) says it is exact in float; other sets' entries
// are left unwritten (the kernels do not read them). pt_off: n_sets+1 point
// offsets, or nullptr for one set. pk (one set only, or nullptr): when the
// set is packable, its packed copy is written instead of the float copy.
__global__ __launch_bounds__(256) void k_points_f32(const double* __restrict__ xyz,
                                                    const int64_t* __restrict__ pt_off, uint32_t n_sets,
                                                    int64_t total, const int32_t* __restrict__ grid,
                                                    F4* __restrict__ out, uint32_t* __restrict__ pk) {
    if (pk) {
        const SetGrid g{grid[0], grid[1]};
        if (set_packable(g)) {
            const double inv = pow2i(-g.gmin);
            for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
                 i += (int64_t)gridDim.x * blockDim.x)
                pk[i] = pack_point(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], inv);
            return;
        }
    }
    // Each workgroup takes one contiguous range of points, so a thread's set
    // changes only at set boundaries (looked up then, from the previous set
    // on); the points of sets that are not exact in float are skipped whole.
    const int64_t per = ((total + gridDim.x - 1) / gridDim.x + 255) / 256 * 256;
    const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(total, b0 + per);
    int64_t cur_end = pt_off ? 0 : total;
    uint32_t set = 0;
    bool exact = pt_off ? false : set_f32_exact(SetGrid{grid[0], grid[1]});
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 256) {
        if (i >= cur_end) {  // pt_off != nullptr here: last s with pt_off[s] <= i
            uint32_t lo = set, hi = n_sets;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (pt_off[mid] <= i) lo = mid;
                else hi = mid;
            }
            set = lo;
            cur_end = pt_off[set + 1];
            exact = set_f32_exact(SetGrid{grid[2 * set], grid[2 * set + 1]});
        }
        if (!exact) {
            // the last of this thread's points that precedes the set's end
            i += (cur_end - 1 - i) / 256 * 256;
            continue;
        }
        out[i] = F4{(float)xyz[3 * i], (float)xyz[3 * i + 1], (float)xyz[3 * i + 2], 0.0f};
    }
}

void enqueue_points_f32(const double* xyz, const int64_t* pt_off, uint32_t n_sets, int64_t total,
                        const int32_t* grid, F4* out, hipStream_t s, uint32_t* pk) {
    if (total <= 0) return;
    const int cus = device_cu_count();
    const unsigned g = (unsigned)std::min<int64_t>((total + 255) / 256, (int64_t)cus * 8);
    k_points_f32<<<g, 256, 0, s>>>(xyz, pt_off, n_sets, total, grid, out, pk);
    D3_LAUNCH_CHECK();
}

void enqueue_set_grids(const BatchArgs& a, uint32_t n_sets, int64_t total_pts, int32_t* grid,
                       hipStream_t s, F4* out32) {
    k_grid_init<<<(n_sets + 255) / 256, 256, 0, s>>>(grid, n_sets, a.ctr);
    D3_LAUNCH_CHECK();
    if (total_pts <= 0) return;
    const int cus = device_cu_count();
    const int64_t step = 256 * kGridUnroll;
    const int64_t units = out32 && !a.set_ptr ? total_pts : 3 * total_pts;  // points or doubles
    const unsigned g = (unsigned)std::min<int64_t>((units + step - 1) / step, (int64_t)cus * 8);
    k_grid_points<<<g, 256, 0, s>>>(a, n_sets, total_pts, a.set_ptr ? a.pt_first : a.pt_off, grid,
                                    a.set_ptr ? nullptr : out32);
    D3_LAUNCH_CHECK();
}

// Zeroes the per-chunk counters (a kernel, so the whole sequence can be
// captured into a graph).
__global__ void k_clear_counters(PredCounters* c) {
    if (threadIdx.x < sizeof(PredCounters) / 4) reinterpret_cast<uint32_t*>(c)[threadIdx.x] = 0u;
}

void enqueue_batch(const BatchArgs& base, const BatchShape& sh, hipStream_t s, hipEvent_t* ev,
                   hipEvent_t* ev2) {
    const int cus = device_cu_count();
    if (!sh.counters_cleared) {
        k_clear_counters<<<1, 64, 0, s>>>(base.ctr);
        D3_LAUNCH_CHECK();
    }
    if (ev) D3_HIP(hipEventRecord(ev[0], s));
    if (ev2) D3_HIP(hipEventRecord(ev2[0], s));
    if (sh.n_lds_tasks) {
        BatchArgs a = base;
        a.tasks = sh.lds_tasks;
        a.n_tasks = sh.n_lds_tasks;
        a.chunk = sh.lds_chunk;
        const unsigned g = (unsigned)std::min<uint64_t>(sh.n_lds_tasks, (uint64_t)cus * sh.lds_blocks_per_cu);
        if (sh.packed) {
            if (a.stats) pred_batch_packed<true><<<g, kLdsThreads, 0, s>>>(a);
            else pred_batch_packed<false><<<g, kLdsThreads, 0, s>>>(a);
        } else {
            if (a.stats) pred_batch_lds<true><<<g, kLdsThreads, 0, s>>>(a);
            else pred_batch_lds<false><<<g, kLdsThreads, 0, s>>>(a);
        }
        D3_LAUNCH_CHECK();
    }
    if (sh.n_global_tasks) {
        BatchArgs a = base;
        a.tasks = sh.global_tasks;
        a.n_tasks = sh.n_global_tasks;
        a.chunk = sh.global_chunk;
        a.bins = sh.global_tasks ? sh.global_bins : nullptr;
        const unsigned g =
            (unsigned)std::min<uint64_t>(sh.n_global_tasks, (uint64_t)cus * sh.global_blocks_per_cu);
        if (a.stats) pred_batch_global<true><<<g, kGlobalThreads, 0, s>>>(a);
        else pred_batch_global<false><<<g, kGlobalThreads, 0, s>>>(a);
        D3_LAUNCH_CHECK();
        if (sh.global_packed) {
            if (a.stats) pred_batch_global_packed<true><<<g, kGlobalThreads, 0, s>>>(a);
            else pred_batch_global_packed<false><<<g, kGlobalThreads, 0, s>>>(a);
            D3_LAUNCH_CHECK();
        }
    }
    if (ev2) D3_HIP(hipEventRecord(ev2[1], s));
    if (ev) {
        D3_HIP(hipEventRecord(ev[1], s));
        D3_HIP(hipEventRecord(ev[2], s));
    }
    pred_exact<<<cus * 2, 256, 0, s>>>(base);  // 2 wavefronts per SIMD: its VGPR limit
    D3_LAUNCH_CHECK();
    pred_exact_mid<<<cus * 2, 256, 0, s>>>(base);  // its VGPR limit: 2 wavefronts per SIMD
    D3_LAUNCH_CHECK();
    if (ev) {
        D3_HIP(hipEventRecord(ev[3], s));
        D3_HIP(hipEventRecord(ev[4], s));
    }
    pred_wide<<<cus * 2, kWideBlock, 0, s>>>(base);
    D3_LAUNCH_CHECK();
    if (ev) D3_HIP(hipEventRecord(ev[5], s));
}

// The scratch one batch call uses: the exact tiers' queues (one entry per
// query of a launch chunk), the counters, the set's grid summary, and the
// float or packed copy of the points.
struct PredScratch {
    uint32_t* queue = nullptr;
    uint32_t* wide = nullptr;
    PredCounters* ctr = nullptr;
    int32_t* grid = nullptr;
    F4* pts32 = nullptr;
    uint32_t* pk = nullptr;
    uint32_t chunk = 0;  // queries per launch sequence
};

struct PredRoute {
    bool lds, packed, global_packed;
};

static PredRoute route_for(int64_t n_pts, int force_path) {
    PredRoute r;
    r.lds = force_path == 1 || (force_path == 0 && n_pts <= kLdsPts);
    r.packed = force_path == 0 && n_pts > kLdsPts && n_pts <= kPackPts;
    r.global_packed = force_path == 0 && n_pts > kPackPts;
    return r;
}

// One point set: the LDS kernel when the set fits; else, up to kPackPts
// points, the packed LDS kernel and the global one, of which the set's
// summary runs one (packable sets: the packed kernel); else the global
// kernel and its packed variant, again one of them by the summary.
// The tasks are implicit equal chunks. Queries are processed in launch
// sequences of at most b.chunk (which bounds the queues).
static void run_predicates(const PredScratch& b, const PredRoute& r, const double* xyz, int64_t n_pts,
                           const uint8_t* op, const int32_t* idx, int64_t n_q, int8_t* out, hipStream_t s,
                           unsigned long long* stats_dev, hipEvent_t* ev, hipEvent_t* ev2) {
    const int cus = device_cu_count();
    bool grids_done = false;
    for (int64_t off = 0; off < n_q; off += b.chunk) {
        const int64_t n = std::min<int64_t>(b.chunk, n_q - off);
        BatchArgs a{};
        a.xyz = xyz;
        a.n_pts = n_pts;
        a.op = op + off;
        a.idx = idx + off * 5;
        a.out = out + off;
        a.n_q = n;
        a.queue = b.queue;
        a.wide = b.wide;
        a.ctr = b.ctr;
        a.stats = stats_dev;
        a.grid = b.grid;
        a.pts32 = b.pts32;
        a.pk = b.pk;
        a.pack_route = (r.packed || r.global_packed) ? 1 : 0;
        BatchShape sh{};
        if (!grids_done) {
            // the set's grid summary, and this chunk's counters, in one prepass
            enqueue_set_grids(a, 1, n_pts, b.grid, s);
            if (b.pts32) enqueue_points_f32(xyz, nullptr, 1, n_pts, b.grid, b.pts32, s, b.pk);
            grids_done = true;
            sh.counters_cleared = true;
        }
        if (r.lds || r.packed) {
            // about two tasks per CU, at least 8 tiles each
            const int64_t per = std::max<int64_t>(8 * kLdsThreads, (n + 2 * cus - 1) / (2 * cus));
            sh.lds_chunk = (uint32_t)std::min<int64_t>(per, n);
            sh.n_lds_tasks = (uint32_t)((n + sh.lds_chunk - 1) / sh.lds_chunk);
            sh.lds_blocks_per_cu = 1;
            sh.packed = r.packed;
        }
        if (!r.lds) {
            // 4 workgroups of 4 wavefronts per CU (4 wavefronts per SIMD, the
            // kernel's occupancy), about two tasks each, tasks of 4..64 tiles
            // (measured on 2^24 lattice queries: 32 tiles x 4 per CU is 1.3%
            // faster than 16 tiles x 8 per CU)
            sh.global_blocks_per_cu = 4;
            const int64_t tiles = (n + kGlobalThreads - 1) / kGlobalThreads;
            const int64_t per = std::min<int64_t>(64, std::max<int64_t>(4, tiles / (2 * 4 * (int64_t)cus)));
            sh.global_chunk = (uint32_t)(per * kGlobalThreads);
            sh.n_global_tasks = (uint32_t)((n + sh.global_chunk - 1) / sh.global_chunk);
            sh.global_packed = r.global_packed;
        }
        enqueue_batch(a, sh, s, ev, ev2);
    }
}

// Calls of at most this many queries (without statistics, timing events or
// a forced path) run the one-pass kernel: no grid-summary prepass and no
// separate exact launches, whose fixed cost (about 25 us a call) outweighs
// the pipeline's faster fast path for small calls. Measured per call, the
// pipeline -> the one-pass kernel: on 10^4 random points, 2^12 queries
// 42 -> 21 us, 2^16 50 -> 29 us, 2^18 54 -> 38 us, 2^20 92 -> 64 us; on the
// 8^3 integer lattice (mostly certificate and exact work), 2^12 30 -> 12 us,
// 2^18 44 -> 19 us, but 2^20 46 -> 49 us and 2^21 50 -> 85 us.
constexpr int64_t kSmallBatch = int64_t(1) << 19;

void launch_predicates(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                       int64_t n_q, int8_t* out, hipStream_t s, unsigned long long* stats_dev,
                       hipEvent_t* ev, int force_path, hipEvent_t* ev2) {
    if (n_q <= 0) return;
    if (n_q <= kSmallBatch && !stats_dev && !ev && !ev2 && force_path == 0) {
        Scratch sc(s);
        PredCounters* ctr = sc.get<PredCounters>(1);
        uint32_t* wide = sc.get<uint32_t>((size_t)n_q);
        launch_predicates_small(xyz, n_pts, op, idx, n_q, out, ctr, wide, s);
        return;
    }
    const PredRoute r = route_for(n_pts, force_path);
    PredScratch b;
    b.chunk = (uint32_t)std::min<int64_t>(n_q, kPredChunk);
    Scratch sc(s);
    b.queue = sc.get<uint32_t>(b.chunk);
    b.wide = sc.get<uint32_t>(b.chunk);
    b.ctr = sc.get<PredCounters>(1);
    b.grid = sc.get<int32_t>(2);
    // the global kernel's compact float copy (used only if the set is exact in float)
    b.pts32 = r.lds ? nullptr : sc.get<F4>((size_t)n_pts);
    // the packed copy (written and used only if the set is packable, in place
    // of the float copy): held in LDS by the packed kernel, or gathered from
    // global memory by the global one's packed variant for larger sets
    b.pk = (r.packed || r.global_packed) ? sc.get<uint32_t>((size_t)n_pts) : nullptr;
    run_predicates(b, r, xyz, n_pts, op, idx, n_q, out, s, stats_dev, ev, ev2);
}

void launch_predicates_small(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                             int64_t n_q, int8_t* out, PredCounters* ctr, uint32_t* wide, hipStream_t s) {
    if (n_q <= 0) return;
    BatchArgs a{};
    a.xyz = xyz;
    a.n_pts = n_pts;
    a.op = op;
    a.idx = idx;
    a.out = out;
    a.n_q = n_q;
    a.wide = wide;
    a.ctr = ctr;
    k_clear_counters<<<1, 64, 0, s>>>(ctr);
    D3_LAUNCH_CHECK();
    const unsigned g = (unsigned)std::min<int64_t>((n_q + 255) / 256, (int64_t)device_cu_count() * 8);
    pred_small<<<g, 256, 0, s>>>(a);
    D3_LAUNCH_CHECK();
    const unsigned gw = (unsigned)std::min<int64_t>(n_q, 64);
    pred_wide<<<gw, kWideBlock, 0, s>>>(a);
    D3_LAUNCH_CHECK();
}

// ---------------------------------------------------------------------------
// Allocation-free form (del3d_predicates_noalloc): the same pipeline with its
// scratch carved from one file-scope device array, for callers that cannot
// allocate device memory. Each device holds its own copy of the array until
// the process ends; calls that overlap on one device share it and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// The array's device address on the current device, looked up once per
// device (the array's name in host code is not a device address).
static unsigned char* noalloc_scratch_ptr() {
    static std::mutex mu;
    static std::vector<unsigned char*> ptrs;
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)ptrs.size() <= dev) ptrs.resize(dev + 1, nullptr);
    if (!ptrs[dev]) {
        void* p = nullptr;
        D3_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_error(DEL3D_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
            throw HipError{DEL3D_ERR_DEVICE};
        }
        ptrs[dev] = static_cast<unsigned char*>(p);
    }
    return ptrs[dev];
}

void launch_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                               int64_t n_q, int8_t* out, hipStream_t s) {
    if (n_q <= 0) return;
    unsigned char* base = noalloc_scratch_ptr();
    const PredRoute r = route_for(n_pts, 0);
    PredScratch b;
    b.chunk = (uint32_t)std::min<int64_t>(n_q, kNoAllocChunk);
    size_t at = 0;
    auto carve = [&](size_t bytes) {
        unsigned char* p = base + at;
        at += (bytes + 255) & ~size_t(255);
        return p;
    };
    b.queue = reinterpret_cast<uint32_t*>(carve(sizeof(uint32_t) * kNoAllocChunk));
    b.wide = reinterpret_cast<uint32_t*>(carve(sizeof(uint32_t) * kNoAllocChunk));
    b.ctr = reinterpret_cast<PredCounters*>(carve(sizeof(PredCounters)));
    b.grid = reinterpret_cast<int32_t*>(carve(2 * sizeof(int32_t)));
    b.pts32 = r.lds ? nullptr : reinterpret_cast<F4*>(carve(sizeof(F4) * (size_t)n_pts));
    b.pk = (r.packed || r.global_packed) ? reinterpret_cast<uint32_t*>(carve(sizeof(uint32_t) * (size_t)n_pts))
                                         : nullptr;
    if (at > kNoAllocScratchBytes) {  // n_pts is checked by the caller; this is the layout's own bound
        set_error(DEL3D_ERR_INVALID, "the allocation-free form's scratch (%zu bytes) cannot hold %zu bytes",
                  kNoAllocScratchBytes, at);
        throw HipError{DEL3D_ERR_INVALID};
    }
    run_predicates(b, r, xyz, n_pts, op, idx, n_q, out, s, nullptr, nullptr, nullptr);
}

}  // namespace d3
