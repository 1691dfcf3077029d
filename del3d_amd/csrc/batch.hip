// batch.hip - predicates over many point sets in one launch sequence: plans
// (task tables built once, then enqueue-only runs), graph replay, autotuning.
//
// A plan covers n_sets point sets; set s owns points [pt_off[s], pt_off[s+1])
// and queries [q_off[s], q_off[s+1]) whose indices are local to the set. Its
// queries are cut into tasks: sets of at most lds_max_pts points go to the LDS
// kernel (the whole set is loaded into one CU's LDS per task, so tasks are
// large), larger sets to the global-gather kernel. Tasks are ordered largest
// first, and workgroups take them dynamically, so the tail is short. Task
// lists are split into launch chunks of at most kPlanChunk queries (the exact
// queues' capacity, allocated with the plan: 8 bytes per query of a chunk). Running a plan only enqueues kernels: no allocation, no
// copy, no host synchronization, so a run can be captured into a HIP graph;
// the plan keeps one instantiated graph per (buffers, stream) and replays it.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "common.h"
#include "predicates.h"

struct del3d_pred_plan_s {
    int dev = 0;
    int64_t n_sets = 0, n_pts = 0, n_q = 0;
    std::vector<int64_t> pt_off, q_off;  // host copies
    int32_t knobs[DEL3D_PLAN_KNOBS] = {};
    // device tables and scratch (hipMalloc, owned)
    int64_t* d_pt_off = nullptr;
    int64_t* d_q_off = nullptr;
    d3::PredTask* d_tasks = nullptr;
    size_t d_tasks_cap = 0;
    uint32_t* d_bins = nullptr;  // per chunk, kBins+1 offsets of the XCD bins of its global tasks
    size_t d_bins_cap = 0;
    bool binned = false;
    uint32_t* queue = nullptr;
    uint32_t* wide = nullptr;
    d3::PredCounters* ctr = nullptr;
    int32_t* grid = nullptr;  // per set SetGrid summary, recomputed by every run
    d3::F4* pts32 = nullptr;  // float copy of the points (used for float-exact sets), rewritten by every run
    uint32_t queue_cap = 0;
    struct Chunk {
        uint32_t lds0, n_lds, glob0, n_glob;  // task ranges in d_tasks
        uint32_t bins0;                       // the chunk's bin offsets in d_bins (when binned)
    };
    std::vector<Chunk> chunks;
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* g_key[4] = {};
    hipStream_t g_stream = nullptr;
};

namespace d3 {
namespace {

// Queries per launch sequence of a plan. Each sequence ends with the exact
// tiers' launches and the fast kernel's tail, so fewer, larger ones are
// faster: 2^28 (2 GB of queues) puts the 1.1 G queries of 4096 sets of
// 10^2..2*10^5 points in 5 sequences instead of 17.
constexpr int64_t kPlanChunk = int64_t(1) << 28;

void default_knobs(int32_t* k) {
    k[0] = kLdsPts;             // lds_max_pts: largest set routed to the LDS kernel
    k[1] = 32 * kLdsThreads;    // lds_task_queries
    k[2] = 16 * kGlobalThreads; // global_task_queries
    k[3] = 8;                   // global_blocks_per_cu
    k[4] = 1;                   // global tasks in per-XCD bins (0: one shared list)
}

void drop_graph(del3d_pred_plan_s* p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    std::memset(p->g_key, 0, sizeof p->g_key);
    p->g_stream = nullptr;
}

// Builds the task table from the offsets and the knobs, uploads it.
//
// LDS tasks are ordered largest first. Global tasks stay grouped by set (a
// set's tail task next to its others), and a launch chunk takes whole sets
// while they fit (a set larger than the room left is split at a task
// boundary). With knobs[4] = 1, each chunk's sets are then dealt to kBins
// bins, one per XCD, largest set first to the bin with the fewest queries,
// so that each XCD works through its own sets and keeps them in its L2.
void build_tasks(del3d_pred_plan_s* p, hipStream_t s) {
    const int64_t lds_max = std::min<int64_t>(p->knobs[0], kLdsPts);
    // a task never exceeds a launch chunk
    const int64_t lds_q = std::min<int64_t>(std::max<int64_t>(p->knobs[1], kLdsThreads), kPlanChunk);
    const int64_t glob_q = std::min<int64_t>(std::max<int64_t>(p->knobs[2], kGlobalThreads), kPlanChunk);
    const bool binned = p->knobs[4] != 0;
    std::vector<PredTask> lds;
    struct Group {
        int64_t q;                   // queries of the group
        std::vector<PredTask> tasks;  // tasks of one set, in query order
    };
    std::vector<Group> glob;
    for (int64_t si = 0; si < p->n_sets; ++si) {
        const int64_t n = p->pt_off[si + 1] - p->pt_off[si];
        const int64_t a = p->q_off[si], b = p->q_off[si + 1];
        if (b <= a) continue;
        const bool in_lds = n <= lds_max;
        const int64_t step = in_lds ? lds_q : glob_q;
        if (!in_lds) glob.push_back(Group{b - a, {}});
        for (int64_t q = a; q < b; q += step) {
            PredTask t;
            t.q0 = q;
            t.q1 = std::min(b, q + step);
            t.set = (int32_t)si;
            t.pad = 0;
            (in_lds ? lds : glob.back().tasks).push_back(t);
        }
    }
    auto bigger = [](const PredTask& x, const PredTask& y) {
        return (x.q1 - x.q0) != (y.q1 - y.q0) ? (x.q1 - x.q0) > (y.q1 - y.q0) : x.q0 < y.q0;
    };
    std::stable_sort(lds.begin(), lds.end(), bigger);
    std::stable_sort(glob.begin(), glob.end(), [](const Group& x, const Group& y) { return x.q > y.q; });
    // launch chunks: at most kPlanChunk queries each
    p->chunks.clear();
    std::vector<PredTask> all;
    std::vector<uint32_t> bins;
    size_t il = 0, ig = 0, it = 0;  // next LDS task; next global group and its next task
    while (il < lds.size() || ig < glob.size()) {
        del3d_pred_plan_s::Chunk c{};
        int64_t q = 0;
        std::vector<PredTask> cl;
        std::vector<Group> cg;
        while (il < lds.size() && q + (lds[il].q1 - lds[il].q0) <= kPlanChunk) {
            q += lds[il].q1 - lds[il].q0;
            cl.push_back(lds[il++]);
        }
        while (ig < glob.size()) {
            const auto& gt = glob[ig].tasks;
            Group part{0, {}};
            while (it < gt.size() && q + (gt[it].q1 - gt[it].q0) <= kPlanChunk) {
                q += gt[it].q1 - gt[it].q0;
                part.q += gt[it].q1 - gt[it].q0;
                part.tasks.push_back(gt[it++]);
            }
            if (!part.tasks.empty()) cg.push_back(std::move(part));
            if (it < gt.size()) break;  // the chunk is full
            ++ig;
            it = 0;
        }
        c.lds0 = (uint32_t)all.size();
        c.n_lds = (uint32_t)cl.size();
        all.insert(all.end(), cl.begin(), cl.end());
        c.glob0 = (uint32_t)all.size();
        c.bins0 = (uint32_t)bins.size();
        if (binned) {
            // cg is largest first: each group goes to the least-loaded bin
            std::vector<std::vector<const Group*>> bin(kBins);
            int64_t load[kBins] = {};
            for (const auto& g : cg) {
                const int b = (int)(std::min_element(load, load + kBins) - load);
                load[b] += g.q;
                bin[b].push_back(&g);
            }
            for (uint32_t b = 0; b < kBins; ++b) {
                bins.push_back((uint32_t)(all.size() - c.glob0));
                for (const Group* g : bin[b]) all.insert(all.end(), g->tasks.begin(), g->tasks.end());
            }
            bins.push_back((uint32_t)(all.size() - c.glob0));
        } else {
            for (const auto& g : cg) all.insert(all.end(), g.tasks.begin(), g.tasks.end());
        }
        c.n_glob = (uint32_t)(all.size() - c.glob0);
        p->chunks.push_back(c);
    }
    if (all.size() > p->d_tasks_cap) {
        if (p->d_tasks) D3_HIP(hipFree(p->d_tasks));
        p->d_tasks = nullptr;
        p->d_tasks_cap = 0;
        D3_HIP(hipMalloc((void**)&p->d_tasks, std::max<size_t>(1, all.size()) * sizeof(PredTask)));
        p->d_tasks_cap = all.size();
    }
    if (bins.size() > p->d_bins_cap) {
        if (p->d_bins) D3_HIP(hipFree(p->d_bins));
        p->d_bins = nullptr;
        p->d_bins_cap = 0;
        D3_HIP(hipMalloc((void**)&p->d_bins, bins.size() * sizeof(uint32_t)));
        p->d_bins_cap = bins.size();
    }
    if (!all.empty()) copy_h2d(p->d_tasks, all.data(), all.size() * sizeof(PredTask), s);
    if (!bins.empty()) copy_h2d(p->d_bins, bins.data(), bins.size() * sizeof(uint32_t), s);
    p->binned = binned;
    drop_graph(p);
}

void run_plan(del3d_pred_plan_s* p, const double* xyz, const uint8_t* op, const int32_t* idx,
              int8_t* out, hipStream_t s, unsigned long long* stats, hipEvent_t* ev) {
    {
        BatchArgs g{};
        g.xyz = xyz;
        g.pt_off = p->d_pt_off;
        g.ctr = p->ctr;
        // the set summaries and, in the same pass, the float copy of the points
        enqueue_set_grids(g, (uint32_t)p->n_sets, p->n_pts, p->grid, s, p->pts32);
    }
    bool first_chunk = true;
    for (const auto& c : p->chunks) {
        BatchArgs a{};
        a.xyz = xyz;
        a.n_pts = p->n_pts;
        a.pt_off = p->d_pt_off;
        a.q_off = p->d_q_off;
        a.n_sets = (uint32_t)p->n_sets;
        a.op = op;
        a.idx = idx;
        a.out = out;
        a.n_q = p->n_q;
        a.queue = p->queue;
        a.wide = p->wide;
        a.ctr = p->ctr;
        a.stats = stats;
        a.grid = p->grid;
        a.pts32 = p->pts32;
        BatchShape sh;
        sh.lds_tasks = p->d_tasks + c.lds0;
        sh.n_lds_tasks = c.n_lds;
        sh.lds_chunk = 0;
        sh.lds_blocks_per_cu = 1;
        sh.global_tasks = p->d_tasks + c.glob0;
        sh.n_global_tasks = c.n_glob;
        sh.global_chunk = 0;
        sh.global_blocks_per_cu = (uint32_t)std::max(1, p->knobs[3]);
        sh.global_bins = p->binned ? p->d_bins + c.bins0 : nullptr;
        sh.counters_cleared = first_chunk;  // the grid prepass zeroed them
        first_chunk = false;
        enqueue_batch(a, sh, s, p->chunks.size() == 1 ? ev : nullptr);
    }
}

int check_tables(const int64_t* pt_off, const int64_t* q_off, int64_t n_sets) {
    if (n_sets < 0) return set_error(DEL3D_ERR_INVALID, "negative set count");
    if (n_sets > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "too many sets");
    if (!pt_off || !q_off) return set_error(DEL3D_ERR_INVALID, "null offset table");
    if (pt_off[0] != 0 || q_off[0] != 0) return set_error(DEL3D_ERR_INVALID, "offsets must start at 0");
    for (int64_t s = 0; s < n_sets; ++s) {
        if (pt_off[s + 1] < pt_off[s] || q_off[s + 1] < q_off[s])
            return set_error(DEL3D_ERR_INVALID, "offsets decrease at set %lld", (long long)s);
        if (pt_off[s + 1] - pt_off[s] > 0x7FFFFFFF)
            return set_error(DEL3D_ERR_INVALID, "set %lld has more than 2^31-1 points", (long long)s);
        if (q_off[s + 1] > q_off[s] && pt_off[s + 1] == pt_off[s])
            return set_error(DEL3D_ERR_INVALID, "set %lld has queries but no points", (long long)s);
    }
    if (q_off[n_sets] > (int64_t)0xFFFFFFFF)
        return set_error(DEL3D_ERR_INVALID, "a plan takes at most 2^32-1 queries");
    return DEL3D_OK;
}

int check_run(del3d_pred_plan_s* p, const double* xyz, const uint8_t* op, const int32_t* idx,
              int8_t* out) {
    if (!p) return set_error(DEL3D_ERR_INVALID, "null plan");
    if (p->n_q > 0 && (!xyz || !op || !idx || !out)) return set_error(DEL3D_ERR_INVALID, "null buffer");
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    if (dev != p->dev) return set_error(DEL3D_ERR_INVALID, "plan was created on device %d", p->dev);
    return DEL3D_OK;
}

}  // namespace
}  // namespace d3

using namespace d3;

extern "C" {

int del3d_pred_plan_create(const int64_t* pt_off, const int64_t* q_off, int64_t n_sets,
                           del3d_pred_plan* plan_out) {
    return guarded([&] {
        if (!plan_out) return set_error(DEL3D_ERR_INVALID, "null plan_out");
        *plan_out = nullptr;
        const int rc = check_tables(pt_off, q_off, n_sets);
        if (rc) return rc;
        auto* p = new del3d_pred_plan_s;
        struct Guard {
            del3d_pred_plan_s* p;
            ~Guard() {
                if (p) del3d_pred_plan_destroy(p);
            }
        } guard{p};
        D3_HIP(hipGetDevice(&p->dev));
        p->n_sets = n_sets;
        p->pt_off.assign(pt_off, pt_off + n_sets + 1);
        p->q_off.assign(q_off, q_off + n_sets + 1);
        p->n_pts = pt_off[n_sets];
        p->n_q = q_off[n_sets];
        default_knobs(p->knobs);
        hipStream_t s = thread_stream();
        D3_HIP(hipMalloc((void**)&p->d_pt_off, (n_sets + 1) * 8));
        D3_HIP(hipMalloc((void**)&p->d_q_off, (n_sets + 1) * 8));
        copy_h2d(p->d_pt_off, pt_off, (n_sets + 1) * 8, s);
        copy_h2d(p->d_q_off, q_off, (n_sets + 1) * 8, s);
        p->queue_cap = (uint32_t)std::min<int64_t>(std::max<int64_t>(p->n_q, 1), kPlanChunk);
        D3_HIP(hipMalloc((void**)&p->queue, (size_t)p->queue_cap * 4));
        D3_HIP(hipMalloc((void**)&p->wide, (size_t)p->queue_cap * 4));
        D3_HIP(hipMalloc((void**)&p->ctr, sizeof(PredCounters)));
        D3_HIP(hipMalloc((void**)&p->grid, (size_t)std::max<int64_t>(n_sets, 1) * 8));
        D3_HIP(hipMalloc((void**)&p->pts32, (size_t)std::max<int64_t>(p->n_pts, 1) * sizeof(F4)));
        build_tasks(p, s);
        guard.p = nullptr;
        *plan_out = p;
        return DEL3D_OK;
    });
}

int del3d_pred_plan_set_knobs(del3d_pred_plan p, const int32_t* knobs, int32_t n) {
    return guarded([&] {
        if (!p || (n > 0 && !knobs)) return set_error(DEL3D_ERR_INVALID, "null argument");
        if (n < 0 || n > DEL3D_PLAN_KNOBS) return set_error(DEL3D_ERR_INVALID, "bad knob count");
        for (int i = 0; i < n; ++i)
            if (knobs[i] < 0) return set_error(DEL3D_ERR_INVALID, "knob %d is negative", i);
        for (int i = 0; i < n; ++i) p->knobs[i] = knobs[i];
        build_tasks(p, thread_stream());
        return DEL3D_OK;
    });
}

int del3d_pred_plan_get_knobs(del3d_pred_plan p, int32_t* knobs) {
    return guarded([&] {
        if (!p || !knobs) return set_error(DEL3D_ERR_INVALID, "null argument");
        for (int i = 0; i < DEL3D_PLAN_KNOBS; ++i) knobs[i] = p->knobs[i];
        return DEL3D_OK;
    });
}

int del3d_pred_plan_run(del3d_pred_plan p, const double* xyz, const uint8_t* op, const int32_t* idx,
                        int8_t* sign_out, void* stream) {
    return guarded([&] {
        const int rc = check_run(p, xyz, op, idx, sign_out);
        if (rc) return rc;
        run_plan(p, xyz, op, idx, sign_out, (hipStream_t)stream, nullptr, nullptr);
        return DEL3D_OK;
    });
}

int del3d_pred_plan_run_stats(del3d_pred_plan p, const double* xyz, const uint8_t* op,
                              const int32_t* idx, int8_t* sign_out, void* stream, int64_t* stats_host) {
    return guarded([&] {
        const int rc = check_run(p, xyz, op, idx, sign_out);
        if (rc) return rc;
        if (!stats_host) return set_error(DEL3D_ERR_INVALID, "null stats_host");
        hipStream_t s = (hipStream_t)stream;
        Scratch sc(s);
        unsigned long long* st = sc.get<unsigned long long>(kStats);
        fill_u32(reinterpret_cast<uint32_t*>(st), 0u, 2 * kStats, s);
        run_plan(p, xyz, op, idx, sign_out, s, st, nullptr);
        unsigned long long h[kStats];
        copy_d2h(h, st, sizeof h, s);
        for (int k = 0; k < kStats; ++k) stats_host[k] = (int64_t)h[k];
        return DEL3D_OK;
    });
}

int del3d_pred_plan_run_graph(del3d_pred_plan p, const double* xyz, const uint8_t* op,
                              const int32_t* idx, int8_t* sign_out, void* stream) {
    return guarded([&] {
        const int rc = check_run(p, xyz, op, idx, sign_out);
        if (rc) return rc;
        hipStream_t s = (hipStream_t)stream;
        const void* key[4] = {xyz, op, idx, sign_out};
        if (!p->exec || s != p->g_stream || std::memcmp(key, p->g_key, sizeof key) != 0) {
            drop_graph(p);
            D3_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                run_plan(p, xyz, op, idx, sign_out, s, nullptr, nullptr);
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            D3_HIP(hipStreamEndCapture(s, &p->graph));
            D3_HIP(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
            std::memcpy(p->g_key, key, sizeof key);
            p->g_stream = s;
        }
        D3_HIP(hipGraphLaunch(p->exec, s));
        return DEL3D_OK;
    });
}

int del3d_pred_plan_autotune(del3d_pred_plan p, const double* xyz, const uint8_t* op,
                             const int32_t* idx, int8_t* sign_out, void* stream, float* best_ms) {
    return guarded([&] {
        const int rc = check_run(p, xyz, op, idx, sign_out);
        if (rc) return rc;
        hipStream_t s = (hipStream_t)stream;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        D3_HIP(hipEventCreate(&e0));
        D3_HIP(hipEventCreate(&e1));
        struct Ev {
            hipEvent_t a, b;
            ~Ev() {
                (void)hipEventDestroy(a);
                (void)hipEventDestroy(b);
            }
        } evs{e0, e1};
        int32_t base[DEL3D_PLAN_KNOBS];
        std::memcpy(base, p->knobs, sizeof base);
        // candidates: LDS routing threshold, task sizes, global occupancy
        std::vector<std::vector<int32_t>> cands;
        for (int32_t lmax : {kLdsPts, 2048, 0})
            for (int32_t lq : {8 * kLdsThreads, 32 * kLdsThreads, 128 * kLdsThreads})
                for (int32_t gb : {4, 8}) cands.push_back({lmax, lq, 16 * kGlobalThreads, gb, 1});
        // the global task size and the XCD bins, at the default routing
        for (int32_t gq : {4 * kGlobalThreads, 64 * kGlobalThreads})
            for (int32_t gb : {4, 8}) cands.push_back({kLdsPts, 32 * kLdsThreads, gq, gb, 1});
        for (int32_t gb : {4, 8}) cands.push_back({kLdsPts, 32 * kLdsThreads, 16 * kGlobalThreads, gb, 0});
        float best = 1e30f;
        std::vector<int32_t> pick(base, base + DEL3D_PLAN_KNOBS);
        for (const auto& c : cands) {
            std::memcpy(p->knobs, c.data(), sizeof base);
            build_tasks(p, s);
            run_plan(p, xyz, op, idx, sign_out, s, nullptr, nullptr);  // warm-up
            float t = 1e30f;
            for (int r = 0; r < 3; ++r) {
                D3_HIP(hipEventRecord(e0, s));
                run_plan(p, xyz, op, idx, sign_out, s, nullptr, nullptr);
                D3_HIP(hipEventRecord(e1, s));
                D3_HIP(hipEventSynchronize(e1));
                float ms = 0.0f;
                D3_HIP(hipEventElapsedTime(&ms, e0, e1));
                t = std::min(t, ms);
            }
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        std::memcpy(p->knobs, pick.data(), sizeof base);
        build_tasks(p, s);
        if (best_ms) *best_ms = best;
        return DEL3D_OK;
    });
}

int del3d_pred_plan_destroy(del3d_pred_plan p) {
    if (!p) return DEL3D_OK;
    drop_graph(p);
    (void)hipFree(p->d_pt_off);
    (void)hipFree(p->d_q_off);
    (void)hipFree(p->d_tasks);
    (void)hipFree(p->d_bins);
    (void)hipFree(p->queue);
    (void)hipFree(p->wide);
    (void)hipFree(p->ctr);
    (void)hipFree(p->grid);
    (void)hipFree(p->pts32);
    delete p;
    return DEL3D_OK;
}

int del3d_predicates_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                          const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                          int8_t* sign_out, void* stream) {
    del3d_pred_plan p = nullptr;
    int rc = del3d_pred_plan_create(pt_off, q_off, n_sets, &p);
    if (rc) return rc;
    rc = del3d_pred_plan_run(p, xyz, op, idx, sign_out, stream);
    if (rc == DEL3D_OK) {
        rc = guarded([&] {
            D3_HIP(hipStreamSynchronize((hipStream_t)stream));  // the plan's scratch is freed next
            return DEL3D_OK;
        });
    }
    del3d_pred_plan_destroy(p);
    return rc;
}

}  // extern "C"
