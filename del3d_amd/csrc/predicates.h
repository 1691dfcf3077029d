// predicates.h - shared types of the batched predicate kernels and their
// host launchers (internal to the library).
#pragma once

#include <cstdint>

#include "common.h"

namespace d3 {

constexpr int kLdsThreads = 1024;   // workgroup of the LDS kernel (one per CU)
constexpr int kGlobalThreads = 256; // workgroup of the global-gather kernel
constexpr int kLdsPts = 5632;       // largest point set held in LDS (132 KB)
constexpr int kPackPts = 33792;     // largest packed point set held in LDS (132 KB, 4 bytes a point)
constexpr uint32_t kChunk = 1u << 26;  // queries per launch sequence of coplanar, circumcentre and served batches
// 3D queries per launch sequence of one call (bounds its exact queues, 8 B per
// query): each sequence ends with the exact tiers' launches (about 100 us of
// latency-bound work however few queries they hold) and a grid pass, so a
// large call runs best as few sequences as memory allows.
constexpr uint32_t kPredChunk = 1u << 28;
constexpr int kStats = 6;           // decision counters (see del3d_predicates_stats)
// XCDs of the chip. Workgroups are dispatched to the XCDs round-robin
// (workgroup b runs on XCD b % 8), and each XCD has its own L2, so a plan
// can give each XCD its own task list of whole point sets: the sets one XCD
// works on at a time then stay in that XCD's L2 (bins, below).
constexpr uint32_t kBins = 8;

// Queries [q0, q1) of point set `set`; their indices are local to the set.
struct PredTask {
    int64_t q0, q1;
    int32_t set, pad;
};

struct PredCounters {
    uint32_t queued;        // entries in the exact queue
    uint32_t wide;          // entries in the wide queue
    uint32_t next_task[2];  // task cursors of the global [0] and LDS [1] kernels
    uint32_t bin_next[kBins];  // task cursors of the global kernel's XCD bins
};

// A point stored as four floats (16 bytes, one aligned load): the compact
// copy of a point set whose coordinates are all exactly representable in
// float (set_f32_exact).
struct alignas(16) F4 {
    float x, y, z, w;
};

// Kernel arguments of one launch sequence (device pointers).
struct BatchArgs {
    const double* xyz;       // all points
    int64_t n_pts;           // points of the single set (pt_off == nullptr)
    const int64_t* pt_off;   // n_sets+1 point offsets, or nullptr (one set)
    const int64_t* q_off;    // n_sets+1 query offsets, or nullptr (one set)
    uint32_t n_sets;
    // alternatively, one device pointer and point count per set (sets in
    // separate allocations, as resident point sets of a server are); used
    // when set_ptr != nullptr, in place of xyz + pt_off
    const double* const* set_ptr;
    const int64_t* set_np;
    const int64_t* pt_first;  // with set_ptr: n_sets+1 prefix of set_np (for the grid prepass)
    const uint8_t* op;
    const int32_t* idx;
    int8_t* out;
    int64_t n_q;             // queries of an implicit (single-set) task list
    const PredTask* tasks;   // explicit task table, or nullptr: chunks of `chunk`
    uint32_t n_tasks;
    uint32_t chunk;
    uint32_t* queue;         // exact-tier queue (capacity: the queries of the launch)
    uint32_t* wide;          // wide-tier queue (same capacity)
    PredCounters* ctr;
    unsigned long long* stats;  // 6 counters or nullptr
    const int32_t* grid;        // per set {gmin, mexp} (SetGrid), or nullptr
    // a float copy of the points (indexed like xyz + pt_off), valid for the
    // sets whose summary says every coordinate is exact in float; nullptr
    // when not made (the per-set pointer layout)
    const F4* pts32;
    // the packed copy of a single set (set_packable), and whether the call
    // runs both the packed LDS kernel and the global one, each of which
    // returns at once when the set's summary routes it to the other
    const uint32_t* pk;
    int32_t pack_route;
    // global tasks in XCD bins (nullptr: one shared list): bin b holds tasks
    // [bins[b], bins[b+1]); a workgroup takes tasks from bin blockIdx.x % kBins
    // first, then from the others once its own is empty
    const uint32_t* bins;
};

// How a batch is split between the two fast kernels.
struct BatchShape {
    const PredTask* lds_tasks = nullptr;     // nullptr: implicit chunks
    uint32_t n_lds_tasks = 0, lds_chunk = 0, lds_blocks_per_cu = 1;
    const PredTask* global_tasks = nullptr;
    uint32_t n_global_tasks = 0, global_chunk = 0, global_blocks_per_cu = 8;
    const uint32_t* global_bins = nullptr;  // kBins+1 offsets into global_tasks, or nullptr
    bool counters_cleared = false;  // the grid prepass already zeroed the counters
    bool packed = false;            // the LDS tasks run the packed-point kernel
    bool global_packed = false;     // the global tasks also run the global packed variant
};

// Writes the SetGrid summary of every set of `a` into grid (2 int32 per set,
// n_sets sets; one set when a.pt_off and a.set_ptr are null), on s; also
// zeroes a.ctr when it is set.
// Writes the float copy of the points of the sets that are exact in float
// (per the summaries in grid), for the global kernel's gathers; with pk (one
// set only), the packed copy instead when the set is packable.
void enqueue_points_f32(const double* xyz, const int64_t* pt_off, uint32_t n_sets, int64_t total,
                        const int32_t* grid, F4* out, hipStream_t s, uint32_t* pk = nullptr);
// With out32 (point numbering a.pt_off, n_pts entries), the same pass also
// writes the float copy of every point (read only for sets exact in float).
void enqueue_set_grids(const BatchArgs& a, uint32_t n_sets, int64_t total_pts, int32_t* grid,
                       hipStream_t s, F4* out32 = nullptr);

// Enqueues the whole pipeline of one batch on s (no host synchronization;
// capturable into a graph). ev (6 events) brackets fast / exact / wide.
void enqueue_batch(const BatchArgs& a, const BatchShape& sh, hipStream_t s, hipEvent_t* ev,
                   hipEvent_t* ev2 = nullptr);

// One point set; force_path 0 = choose, 1 = LDS kernel, 2 = global kernel.
// The choice: the LDS kernel up to kLdsPts points; then up to kPackPts the
// packed LDS kernel beside the global one (the set's summary runs one of
// them); the global kernel above.
// ev: 6 events (fast, exact, wide brackets) or nullptr; ev2: 2 events around
// the fast kernels only, or nullptr.
void launch_predicates(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                       int64_t n_q, int8_t* out, hipStream_t s, unsigned long long* stats_dev,
                       hipEvent_t* ev = nullptr, int force_path = 0, hipEvent_t* ev2 = nullptr);

// The allocation-free form: at most kNoAllocPts points, queries in launch
// sequences of kNoAllocChunk, all scratch in one file-scope device array of
// kNoAllocScratchBytes (queues 2 x 4 MiB, float and packed copies 20 B per
// point, counters and the grid summary).
constexpr int64_t kNoAllocPts = int64_t(1) << 20;
constexpr uint32_t kNoAllocChunk = 1u << 20;
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
void launch_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                               int64_t n_q, int8_t* out, hipStream_t s);

// Small batch of 3D queries (op codes DEL3D_OP_*) in one pass, each lane all
// tiers (no set summary, no exact queue); queries wider than 4 limbs go to
// the wide queue `wide` (capacity n_q) and pred_wide. ctr is cleared first.
// Same results as launch_predicates.
void launch_predicates_small(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                             int64_t n_q, int8_t* out, PredCounters* ctr, uint32_t* wide, hipStream_t s);

// Coplanar batch (predicates2.hip): op codes DEL3D_OP2_* + op_base, idx 5
// per query (the op's points in argument order), out = the predicate's
// result. Queries with an op byte below op_base are left untouched (another
// kernel of a mixed batch owns them); an unknown op or an index outside
// [0, n_pts) gives 0.
void launch_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* out, hipStream_t s, uint8_t op_base = 0);

// Circumcentres (circumcenter.hip): cells 4 int32 each, out 3 doubles per
// cell; lazy selects circumcenter_lazy, else the exact circumcenter.
void launch_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n_cells,
                          int lazy, double* out, hipStream_t s);

}  // namespace d3
