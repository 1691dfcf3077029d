// server.hip - a serving loop for predicate and sort jobs (continuous
// batching).
//
// Clients register point sets (uploaded once, kept resident in HBM) and
// submit jobs on them: a batch of predicate queries, or a Hilbert sort of the
// whole set, with host buffers. A scheduler thread admits queued jobs into
// batches and hands each batch to a free slot; each slot has its own thread,
// stream and device arenas, so the copies of one batch overlap the kernels
// and copies of others, and jobs that arrive while batches run are admitted
// as soon as a slot frees up. A predicate batch is one launch sequence of the
// batched predicate kernels over its jobs, each job a set of the batch
// (per-set point pointers), so every job's signs are exactly those of
// del3d_predicates on its set alone; a sort job is a batch of its own, the
// spatial sort of its resident set, exactly del3d_spatial_sort's permutation.
//
// Host buffers: pinned (page-locked) client arrays are copied by DMA directly
// in both directions; pageable ones go through the slot's pinned staging.
//
// Memory: resident sets, slot arenas and running sorts' scratch stay within
// the HBM budget. When an upload or a batch's buffers would exceed it, sets no
// running batch uses are evicted least recently used first, then idle slots'
// buffers are freed; an evicted set is uploaded again from the client's host
// array (which must stay valid while the set is registered) when a job needs
// it. Uploads run outside the scheduler's lock, in the worker of the batch
// that needs the set; jobs of other batches on a set being uploaded wait for
// it. If memory is still short while other batches run, the jobs wait at the
// head of the queue for one to finish; if nothing else runs, a job that still
// does not fit fails with DEL3D_ERR_NOMEM, without affecting other jobs.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "common.h"
#include "predicates.h"

using Clock = std::chrono::steady_clock;

namespace d3 {

void spatial_sort_device(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                         hipStream_t s);

namespace {

struct ResidentSet {
    const double* host = nullptr;
    int64_t n = 0;
    double* dev = nullptr;   // nullptr while evicted
    bool uploading = false;  // dev allocated, its copy running outside the lock
    int refs = 0;            // queued or running jobs (the set cannot be dropped)
    int pinned = 0;          // running jobs and uploads (the set cannot be evicted)
    uint64_t last_use = 0;
};

enum JobKind { kJobPredicates = 0, kJobSort = 1 };

struct Job {
    int kind = kJobPredicates;
    int64_t id = 0, set = 0, n_q = 0;
    const uint8_t* op = nullptr;
    const int32_t* idx = nullptr;
    int8_t* out = nullptr;        // predicate signs
    int32_t* order = nullptr;     // sort result (n of the set)
    Clock::time_point t_submit, t_done;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int rc = DEL3D_OK;
    std::string err;
};

struct Slot {
    hipStream_t s = nullptr;
    hipEvent_t done = nullptr;
    // device arenas
    uint8_t* d_op = nullptr;
    int32_t* d_idx = nullptr;
    int8_t* d_out = nullptr;
    uint32_t* d_queue = nullptr;
    uint32_t* d_wide = nullptr;
    PredCounters* d_ctr = nullptr;
    const double** d_ptr = nullptr;
    int64_t* d_np = nullptr;
    int64_t* d_qoff = nullptr;
    int64_t* d_first = nullptr;
    int32_t* d_grid = nullptr;
    PredTask* d_tasks = nullptr;
    int32_t* d_order = nullptr;  // a sort job's order
    // pinned host tables and staging (staging only for pageable client arrays)
    const double** h_ptr = nullptr;
    int64_t* h_np = nullptr;
    int64_t* h_qoff = nullptr;
    int64_t* h_first = nullptr;
    PredTask* h_tasks = nullptr;
    uint8_t* h_stage = nullptr;
    int64_t stage_bytes = 0;
    int64_t cap_q = 0, cap_jobs = 0, cap_tasks = 0, cap_order = 0;
    int64_t bytes = 0;       // device bytes held by the arenas
    int64_t sort_bytes = 0;  // budget reserved for a running sort's scratch
    // scratch blocks the slot's stream keeps cached after its sorts (the next
    // sort reuses them without hipMalloc); counted in the budget, freed when
    // memory is short and the slot is idle
    int64_t cache_bytes = 0;
    bool busy = false;
    int64_t q_run = 0, n_tasks = 0, n_lds = 0;  // the batch in flight
    std::vector<std::shared_ptr<Job>> jobs;
    std::vector<int64_t> uploads;  // sets this batch uploads before it runs
    // the slot's worker thread runs each batch handed to it
    std::thread worker;
    std::condition_variable cv;
    bool has_work = false, quit = false;
};

}  // namespace
}  // namespace d3

using namespace d3;

struct del3d_server_s {
    int dev = 0;
    int64_t budget = 0, used = 0, max_batch_q = 0;
    std::mutex mu;
    std::condition_variable cv_sched, cv_done;
    bool stop = false;
    std::map<int64_t, ResidentSet> sets;
    std::map<int64_t, std::shared_ptr<Job>> jobs;
    std::deque<std::shared_ptr<Job>> queue;
    std::vector<std::unique_ptr<Slot>> slots;
    int64_t next_set = 1, next_job = 1;
    uint64_t tick = 0;
    bool blocked = false;  // the queue head waits for memory or an upload
    int64_t stat_evictions = 0, stat_uploads = 0, stat_batches = 0, stat_done = 0;
    std::thread sched;
};

namespace d3 {
namespace {

constexpr int64_t kTaskQ = 16 * kGlobalThreads;  // queries per task (both kernels)

int64_t tasks_for(int64_t n_q) { return (n_q + kTaskQ - 1) / kTaskQ; }

// Device scratch budgeted for a spatial sort of n points: two record
// buffers (56 B per point), the shuffle tables (about 32 B per point) and the
// round tables, plus the scratch cache's block rounding (up to 2x below
// 1 MiB per block, 1 MiB per block above).
int64_t sort_scratch_bytes(int64_t n) {
    const int64_t m = int64_t(1) << 20;
    return (n <= m ? 192 * n : 96 * n + 96 * m) + (int64_t(256) << 10);
}

// Frees every arena of a slot (device and pinned) and returns its bytes.
void slot_release(del3d_server_s* sv, Slot& sl) {
    for (void* p : {(void*)sl.d_op, (void*)sl.d_idx, (void*)sl.d_out, (void*)sl.d_queue,
                    (void*)sl.d_wide, (void*)sl.d_ctr, (void*)sl.d_ptr, (void*)sl.d_order})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)sl.h_ptr, (void*)sl.h_stage})
        if (p) (void)hipHostFree(p);
    sl.d_op = nullptr;
    sl.d_idx = nullptr;
    sl.d_out = nullptr;
    sl.d_queue = sl.d_wide = nullptr;
    sl.d_ctr = nullptr;
    sl.d_ptr = nullptr;
    sl.d_np = sl.d_qoff = sl.d_first = nullptr;
    sl.d_grid = nullptr;
    sl.d_tasks = nullptr;
    sl.d_order = nullptr;
    sl.h_ptr = nullptr;
    sl.h_np = sl.h_qoff = sl.h_first = nullptr;
    sl.h_tasks = nullptr;
    sl.h_stage = nullptr;
    sl.stage_bytes = 0;
    sv->used -= sl.bytes;
    sl.bytes = 0;
    sl.cap_q = sl.cap_jobs = sl.cap_tasks = sl.cap_order = 0;
}

// set pointers [j], point counts [j], query offsets [j+1], first points [j+1],
// grid summaries [2j int32], tasks [t]
size_t table_bytes(int64_t j, int64_t t) {
    return (size_t)(8 * j + 8 * j + 8 * (j + 1) + 8 * (j + 1) + 8 * j + t * (int64_t)sizeof(PredTask));
}

// Device bytes of a slot sized for q queries, j jobs, t tasks and a sort of o.
int64_t slot_bytes(int64_t q, int64_t j, int64_t t, int64_t o) {
    return q * (1 + 20 + 1 + 4 + 4) + (int64_t)sizeof(PredCounters) + (int64_t)table_bytes(j, t) + 4 * o;
}

// Frees the scratch blocks an idle slot's stream keeps cached (hipFree waits
// for the device) and returns their bytes. Caller holds sv->mu.
void slot_release_cache(del3d_server_s* sv, Slot& sl) {
    if (sl.cache_bytes == 0) return;
    release_stream_cache(sl.s);
    sv->used -= sl.cache_bytes;
    sl.cache_bytes = 0;
}

// Evicts resident sets no running job or upload uses, least recently used
// first, then frees the arenas of idle slots (a slot being filled
// reallocates its own), until `need` more bytes fit the budget. Returns
// whether they do. Caller holds sv->mu.
bool make_room(del3d_server_s* sv, int64_t need) {
    while (sv->used + need > sv->budget) {
        int64_t victim = -1;
        uint64_t oldest = ~0ull;
        for (auto& kv : sv->sets) {
            const ResidentSet& r = kv.second;
            if (r.dev && r.pinned == 0 && !r.uploading && r.last_use < oldest) {
                oldest = r.last_use;
                victim = kv.first;
            }
        }
        if (victim < 0) {
            Slot* idle = nullptr;
            for (auto& sl : sv->slots)
                if (!sl->busy && (sl->bytes > 0 || sl->cache_bytes > 0)) idle = sl.get();
            if (!idle) return false;
            slot_release(sv, *idle);
            slot_release_cache(sv, *idle);
            continue;
        }
        ResidentSet& r = sv->sets[victim];
        (void)hipFree(r.dev);
        r.dev = nullptr;
        sv->used -= r.n * 24;
        ++sv->stat_evictions;
    }
    return true;
}

// Allocates a registered set's device copy and marks it for upload (the
// caller copies it outside the lock, then clears `uploading`). Caller holds
// sv->mu. Throws HipError{DEL3D_ERR_NOMEM} if it cannot fit.
void allocate_set(del3d_server_s* sv, ResidentSet& r) {
    const int64_t bytes = r.n * 24;
    if (!make_room(sv, bytes))
        throw HipError{set_error(DEL3D_ERR_NOMEM, "point set of %lld bytes does not fit the HBM budget",
                                 (long long)bytes)};
    D3_HIP(hipMalloc((void**)&r.dev, (size_t)std::max<int64_t>(bytes, 24)));
    sv->used += bytes;
    ++sv->stat_uploads;
    r.uploading = true;
}

// Host -> device copy on s: DMA straight from pinned memory, through the
// stream's pinned staging otherwise. Returns when enqueued (pinned) or done
// (staged).
void h2d(void* dst, const void* src, size_t bytes, bool pinned, hipStream_t s) {
    if (!bytes) return;
    if (pinned) D3_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    else copy_h2d(dst, src, bytes, s);
}

// Copies a set's points to its device copy on s and waits for the copy.
void upload_set(const double* host, double* dev, int64_t n, hipStream_t s) {
    h2d(dev, host, (size_t)n * 24, host_is_pinned(host), s);
    D3_HIP(hipStreamSynchronize(s));
}

// Grows a slot's arenas to hold q queries, j jobs, t tasks and a sort order
// of o entries. Caller holds sv->mu.
void slot_reserve(del3d_server_s* sv, Slot& sl, int64_t q, int64_t j, int64_t t, int64_t o) {
    if (q <= sl.cap_q && j <= sl.cap_jobs && t <= sl.cap_tasks && o <= sl.cap_order) return;
    // Reallocating frees the old arenas, and hipFree waits for the whole
    // device, stalling the other slots: grow each capacity that must grow to
    // at least twice its old value, so a slot reallocates a few times only.
    auto grow = [](int64_t want, int64_t cap, int64_t limit) {
        if (want <= cap) return cap;
        return std::max(want, std::min(2 * cap, limit));
    };
    q = grow(q, sl.cap_q, std::max(q, sv->max_batch_q));
    j = grow(j, sl.cap_jobs, std::max<int64_t>(j, 4096));
    t = grow(t, sl.cap_tasks, std::max(t, tasks_for(sv->max_batch_q) + 4096));
    o = grow(o, sl.cap_order, std::max<int64_t>(o, int64_t(1) << 31));
    const int64_t want = slot_bytes(q, j, t, o);
    slot_release(sv, sl);
    if (!make_room(sv, want))
        throw HipError{set_error(DEL3D_ERR_NOMEM, "batch buffers of %lld bytes do not fit the HBM budget",
                                 (long long)want)};
    sv->used += want;
    sl.bytes = want;
    if (q > 0) {
        D3_HIP(hipMalloc((void**)&sl.d_op, (size_t)q));
        D3_HIP(hipMalloc((void**)&sl.d_idx, (size_t)q * 20));
        D3_HIP(hipMalloc((void**)&sl.d_out, (size_t)q));
        D3_HIP(hipMalloc((void**)&sl.d_queue, (size_t)q * 4));
        D3_HIP(hipMalloc((void**)&sl.d_wide, (size_t)q * 4));
    }
    if (o > 0) D3_HIP(hipMalloc((void**)&sl.d_order, (size_t)o * 4));
    D3_HIP(hipMalloc((void**)&sl.d_ctr, sizeof(PredCounters)));
    // the batch tables live in one block (one copy per batch):
    // set pointers [j], point counts [j], query offsets [j+1], tasks [t]
    const size_t tab = table_bytes(std::max<int64_t>(j, 1), t);
    const int64_t jj = std::max<int64_t>(j, 1);
    char* dt = nullptr;
    D3_HIP(hipMalloc((void**)&dt, tab));
    sl.d_ptr = reinterpret_cast<const double**>(dt);
    sl.d_np = reinterpret_cast<int64_t*>(dt + 8 * jj);
    sl.d_qoff = reinterpret_cast<int64_t*>(dt + 16 * jj);
    sl.d_first = reinterpret_cast<int64_t*>(dt + 16 * jj + 8 * (jj + 1));
    sl.d_grid = reinterpret_cast<int32_t*>(dt + 16 * jj + 16 * (jj + 1));
    sl.d_tasks = reinterpret_cast<PredTask*>(dt + 24 * jj + 16 * (jj + 1));
    char* ht = nullptr;
    D3_HIP(hipHostMalloc((void**)&ht, tab, hipHostMallocDefault));
    sl.h_ptr = reinterpret_cast<const double**>(ht);
    sl.h_np = reinterpret_cast<int64_t*>(ht + 8 * jj);
    sl.h_qoff = reinterpret_cast<int64_t*>(ht + 16 * jj);
    sl.h_first = reinterpret_cast<int64_t*>(ht + 16 * jj + 8 * (jj + 1));
    sl.h_tasks = reinterpret_cast<PredTask*>(ht + 24 * jj + 16 * (jj + 1));
    sl.cap_q = q;
    sl.cap_jobs = j;
    sl.cap_tasks = t;
    sl.cap_order = o;
}

// The slot's pinned staging, grown to `bytes` (host memory, outside the HBM
// budget). Only the slot's worker uses it.
uint8_t* slot_stage(Slot& sl, int64_t bytes) {
    if (bytes > sl.stage_bytes) {
        if (sl.h_stage) D3_HIP(hipHostFree(sl.h_stage));
        sl.h_stage = nullptr;
        sl.stage_bytes = 0;
        D3_HIP(hipHostMalloc((void**)&sl.h_stage, (size_t)bytes, hipHostMallocDefault));
        sl.stage_bytes = bytes;
    }
    return sl.h_stage;
}


void finish_job(del3d_server_s* sv, const std::shared_ptr<Job>& j, int rc, const char* err) {
    j->t_done = Clock::now();
    j->rc = rc;
    if (err) j->err = err;
    auto it = sv->sets.find(j->set);
    if (it != sv->sets.end()) {
        --it->second.refs;
        if (j->state == 1) --it->second.pinned;
    }
    j->state = rc == DEL3D_OK ? 2 : 3;
    ++sv->stat_done;
}

bool other_slot_busy(del3d_server_s* sv, const Slot* me) {
    for (auto& s : sv->slots)
        if (s.get() != me && s->busy) return true;
    return false;
}

__global__ __launch_bounds__(256) void k_iota(int32_t* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = (int32_t)i;
}

// The predicate batch of a slot: the jobs' queries in, the kernels, the
// signs out (each job's signs straight into its array when that is pinned).
void run_predicate_batch(Slot* sl) {
    const int64_t q_run = sl->q_run, t = sl->n_tasks, n_lds = sl->n_lds;
    const size_t nj = sl->jobs.size();
    hipStream_t s = sl->s;
    int64_t off = 0;
    for (auto& j : sl->jobs) {
        h2d(sl->d_op + off, j->op, (size_t)j->n_q, host_is_pinned(j->op), s);
        h2d(sl->d_idx + 5 * off, j->idx, (size_t)j->n_q * 20, host_is_pinned(j->idx), s);
        off += j->n_q;
    }
    // the tables were written for cap_jobs slots: copy the block up to the tasks used
    const size_t tab_used = reinterpret_cast<char*>(sl->h_tasks + t) - reinterpret_cast<char*>(sl->h_ptr);
    D3_HIP(hipMemcpyAsync((void*)sl->d_ptr, sl->h_ptr, tab_used, hipMemcpyHostToDevice, s));
    BatchArgs a{};
    a.set_ptr = sl->d_ptr;
    a.set_np = sl->d_np;
    a.pt_first = sl->d_first;
    a.q_off = sl->d_qoff;
    a.grid = sl->d_grid;
    a.n_sets = (uint32_t)nj;
    a.op = sl->d_op;
    a.idx = sl->d_idx;
    a.out = sl->d_out;
    a.n_q = q_run;
    a.queue = sl->d_queue;
    a.wide = sl->d_wide;
    a.ctr = sl->d_ctr;
    enqueue_set_grids(a, (uint32_t)nj, sl->h_first[nj], sl->d_grid, s);  // also zeroes a.ctr
    BatchShape sh;
    sh.lds_tasks = sl->d_tasks;
    sh.n_lds_tasks = (uint32_t)n_lds;
    sh.global_tasks = sl->d_tasks + n_lds;
    sh.n_global_tasks = (uint32_t)(t - n_lds);
    sh.counters_cleared = true;
    enqueue_batch(a, sh, s, nullptr);
    // signs: pinned arrays by DMA, the others through the staging
    bool staged = false;
    off = 0;
    for (auto& j : sl->jobs) {
        if (host_is_pinned(j->out)) {
            D3_HIP(hipMemcpyAsync(j->out, sl->d_out + off, (size_t)j->n_q, hipMemcpyDeviceToHost, s));
        } else {
            staged = true;
        }
        off += j->n_q;
    }
    uint8_t* st = staged ? slot_stage(*sl, q_run) : nullptr;
    if (staged) D3_HIP(hipMemcpyAsync(st, sl->d_out, (size_t)q_run, hipMemcpyDeviceToHost, s));
    D3_HIP(hipEventRecord(sl->done, s));
    D3_HIP(hipEventSynchronize(sl->done));
    if (staged) {
        off = 0;
        for (auto& j : sl->jobs) {
            if (!host_is_pinned(j->out)) std::memcpy(j->out, st + off, (size_t)j->n_q);
            off += j->n_q;
        }
    }
}

// A sort job: the spatial sort of its resident set (order 0..n-1), the
// permutation out to the client's array.
void run_sort_job(Slot* sl, const double* dev_pts, int64_t n) {
    const std::shared_ptr<Job>& j = sl->jobs[0];
    hipStream_t s = sl->s;
    const int cus = device_cu_count();
    const unsigned g = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, (int64_t)cus * 8));
    k_iota<<<g, 256, 0, s>>>(sl->d_order, n);
    D3_LAUNCH_CHECK();
    spatial_sort_device(dev_pts, n, sl->d_order, n, s);
    if (host_is_pinned(j->order)) {
        D3_HIP(hipMemcpyAsync(j->order, sl->d_order, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        D3_HIP(hipStreamSynchronize(s));
    } else {
        copy_d2h(j->order, sl->d_order, (size_t)n * 4, s);
    }
    // the sort's scratch stays cached on this stream for the slot's next
    // sort; slot_run moves its bytes from the job's reservation to the slot
}

// Runs a slot's batch: uploads the sets it needs (outside the lock), then
// the predicate batch or the sort job, then completes the jobs.
void slot_run(del3d_server_s* sv, Slot* sl) {
    int rc = guarded([&] {
        for (int64_t id : sl->uploads) {
            const double* host = nullptr;
            double* dev = nullptr;
            int64_t n = 0;
            {
                std::lock_guard<std::mutex> lock(sv->mu);
                const ResidentSet& r = sv->sets[id];
                host = r.host;
                dev = r.dev;
                n = r.n;
            }
            upload_set(host, dev, n, sl->s);
            std::lock_guard<std::mutex> lock(sv->mu);
            ResidentSet& r = sv->sets[id];
            r.uploading = false;
            --r.pinned;  // the upload's pin; the jobs hold their own
            sv->blocked = false;
            sv->cv_sched.notify_all();
        }
        if (sl->jobs[0]->kind == kJobSort) {
            const double* dev = nullptr;
            int64_t n = 0;
            {
                std::lock_guard<std::mutex> lock(sv->mu);
                const ResidentSet& r = sv->sets[sl->jobs[0]->set];
                dev = r.dev;
                n = r.n;
            }
            run_sort_job(sl, dev, n);
        } else {
            run_predicate_batch(sl);
        }
        return DEL3D_OK;
    });
    std::string err = rc ? last_error_slot() : std::string();
    if (rc) (void)hipStreamSynchronize(sl->s);  // let whatever was enqueued drain
    const bool was_sort = sl->jobs[0]->kind == kJobSort;
    int64_t cached = 0;
    if (was_sort) (void)guarded([&] {
        cached = stream_cache_bytes(sl->s);
        return DEL3D_OK;
    });
    std::lock_guard<std::mutex> lock(sv->mu);
    if (was_sort) {  // the job's reservation ends; the blocks its stream keeps stay counted
        sv->used += cached - sl->cache_bytes;
        sl->cache_bytes = cached;
    }
    if (rc) {
        // a failed upload leaves the set unusable: drop its device copy
        for (int64_t id : sl->uploads) {
            ResidentSet& r = sv->sets[id];
            if (r.uploading) {
                r.uploading = false;
                --r.pinned;
                (void)hipFree(r.dev);
                r.dev = nullptr;
                sv->used -= r.n * 24;
            }
        }
    }
    sl->uploads.clear();
    for (auto& j : sl->jobs) finish_job(sv, j, rc, rc ? err.c_str() : nullptr);
    sl->jobs.clear();
    sv->used -= sl->sort_bytes;
    sl->sort_bytes = 0;
    sl->busy = false;
    sv->blocked = false;  // memory may have been freed
    sv->cv_done.notify_all();
    sv->cv_sched.notify_all();
}

// Builds and launches one batch from the queue into a free slot: one sort
// job alone, or predicate jobs up to the batch size (stopping at a sort
// job). Caller holds the lock; returns false if there was nothing to launch.
bool launch_batch(del3d_server_s* sv, Slot& sl, std::unique_lock<std::mutex>& lock) {
    (void)lock;
    std::vector<std::shared_ptr<Job>> take;
    int64_t q = 0;
    if (!sv->queue.empty() && sv->queue.front()->kind == kJobSort) {
        take.push_back(sv->queue.front());
        sv->queue.pop_front();
    }
    while (take.empty() || take[0]->kind == kJobPredicates) {
        if (sv->queue.empty() || sv->queue.front()->kind != kJobPredicates) break;
        auto& j = sv->queue.front();
        if (!take.empty() && q + j->n_q > sv->max_batch_q) break;
        q += j->n_q;
        take.push_back(j);
        sv->queue.pop_front();
    }
    if (take.empty()) return false;
    // Jobs go back to the head of the queue (in order) when their set is
    // being uploaded by another batch, or when memory is short while other
    // batches run; when nothing else runs, a job that still does not fit
    // fails alone.
    auto defer = [&](std::vector<std::shared_ptr<Job>>& js, size_t from) {
        for (size_t k = js.size(); k-- > from;) {
            if (js[k]->state == 1) {
                --sv->sets[js[k]->set].pinned;
                js[k]->state = 0;
            }
            sv->queue.push_front(js[k]);
        }
        js.resize(from);
        sv->blocked = true;
    };
    sl.uploads.clear();
    std::vector<std::shared_ptr<Job>> run;
    int64_t q_run = 0, t_run = 0;
    for (size_t k = 0; k < take.size(); ++k) {
        auto& j = take[k];
        ResidentSet& r = sv->sets[j->set];
        if (r.uploading && std::find(sl.uploads.begin(), sl.uploads.end(), j->set) == sl.uploads.end()) {
            defer(take, k);  // another batch or a client is uploading it
            break;
        }
        int rc = DEL3D_OK;
        if (!r.dev) {
            rc = guarded([&] {
                allocate_set(sv, r);
                return DEL3D_OK;
            });
            if (rc == DEL3D_OK) {
                ++r.pinned;  // the upload's pin
                sl.uploads.push_back(j->set);
            }
        }
        if (rc == DEL3D_ERR_NOMEM && other_slot_busy(sv, &sl)) {
            defer(take, k);
            break;
        }
        if (rc != DEL3D_OK) {
            finish_job(sv, j, rc, last_error_slot().c_str());
            continue;
        }
        r.last_use = ++sv->tick;
        ++r.pinned;
        j->state = 1;
        run.push_back(j);
        q_run += j->n_q;
        t_run += tasks_for(j->n_q);
    }
    auto undo_uploads = [&]() {  // sets allocated (and sort scratch reserved) for jobs that do not run now
        sv->used -= sl.sort_bytes;
        sl.sort_bytes = 0;
        for (int64_t id : sl.uploads) {
            ResidentSet& r = sv->sets[id];
            r.uploading = false;
            --r.pinned;
            (void)hipFree(r.dev);
            r.dev = nullptr;
            sv->used -= r.n * 24;
        }
        sl.uploads.clear();
    };
    if (run.empty()) {
        undo_uploads();
        sv->cv_done.notify_all();
        return true;
    }
    const bool sort = run[0]->kind == kJobSort;
    const int64_t n_sort = sort ? sv->sets[run[0]->set].n : 0;
    int rc = guarded([&] {
        if (sort) {
            // the scratch first: make_room may free idle slots' arenas, this
            // one's included, which slot_reserve then allocates
            // the blocks the slot's stream already caches serve this sort
            // too: reserve only what may come on top of them
            const int64_t need = std::max<int64_t>(0, sort_scratch_bytes(n_sort) - sl.cache_bytes);
            if (!make_room(sv, need))
                throw HipError{set_error(DEL3D_ERR_NOMEM, "a sort of %lld points needs %lld bytes of scratch "
                                         "beyond the HBM budget", (long long)n_sort, (long long)need)};
            sv->used += need;
            sl.sort_bytes = need;
            slot_reserve(sv, sl, 0, 1, 0, n_sort);
        } else {
            slot_reserve(sv, sl, q_run, (int64_t)run.size(), t_run, 0);
        }
        return DEL3D_OK;
    });
    if (rc == DEL3D_ERR_NOMEM && other_slot_busy(sv, &sl)) {
        undo_uploads();
        defer(run, 0);
        return true;
    }
    if (rc == DEL3D_ERR_NOMEM && run.size() > 1) {
        // nothing else runs: try the first job alone, the others wait
        const int64_t first_set = run[0]->set;
        defer(run, 1);
        sv->blocked = false;
        for (size_t k = 0; k < sl.uploads.size();) {  // keep only the first job's upload
            if (sl.uploads[k] == first_set) {
                ++k;
                continue;
            }
            ResidentSet& r = sv->sets[sl.uploads[k]];
            r.uploading = false;
            --r.pinned;
            (void)hipFree(r.dev);
            r.dev = nullptr;
            sv->used -= r.n * 24;
            sl.uploads.erase(sl.uploads.begin() + k);
        }
        q_run = run[0]->n_q;
        t_run = tasks_for(q_run);
        rc = guarded([&] {
            slot_reserve(sv, sl, q_run, 1, t_run, 0);
            return DEL3D_OK;
        });
    }
    if (rc != DEL3D_OK) {
        undo_uploads();
        for (auto& j : run) finish_job(sv, j, rc, last_error_slot().c_str());
        sv->cv_done.notify_all();
        return true;
    }
    int64_t t = 0, n_lds = 0;
    if (!sort) {
        // host tables (the set pointers are stable while pinned)
        int64_t off = 0;
        sl.h_qoff[0] = 0;
        sl.h_first[0] = 0;
        for (size_t k = 0; k < run.size(); ++k) {
            const Job& j = *run[k];
            const ResidentSet& r = sv->sets[j.set];
            sl.h_ptr[k] = r.dev;
            sl.h_np[k] = r.n;
            sl.h_qoff[k + 1] = off + j.n_q;
            sl.h_first[k + 1] = sl.h_first[k] + r.n;
            for (int64_t a = 0; a < j.n_q; a += kTaskQ) {
                PredTask tk;
                tk.q0 = off + a;
                tk.q1 = off + std::min(j.n_q, a + kTaskQ);
                tk.set = (int32_t)k;
                tk.pad = r.n <= kLdsPts ? 1 : 0;  // kernel class
                sl.h_tasks[t++] = tk;
            }
            off += j.n_q;
        }
        // LDS-class tasks first, then global ones (stable)
        std::stable_partition(sl.h_tasks, sl.h_tasks + t, [](const PredTask& x) { return x.pad == 1; });
        n_lds = std::count_if(sl.h_tasks, sl.h_tasks + t, [](const PredTask& x) { return x.pad == 1; });
        for (int64_t i = 0; i < t; ++i) sl.h_tasks[i].pad = 0;
    }
    sl.jobs = run;
    sl.busy = true;
    sl.q_run = q_run;
    sl.n_tasks = t;
    sl.n_lds = n_lds;
    ++sv->stat_batches;
    // the slot's own thread uploads, runs and completes the batch, so the
    // batches of different slots move their data concurrently
    sl.has_work = true;
    sl.cv.notify_one();
    return true;
}

void slot_worker(del3d_server_s* sv, Slot* sl) {
    (void)hipSetDevice(sv->dev);
    for (;;) {
        {
            std::unique_lock<std::mutex> lock(sv->mu);
            sl->cv.wait(lock, [&] { return sl->has_work || sl->quit; });
            if (!sl->has_work) return;
            sl->has_work = false;
        }
        slot_run(sv, sl);
    }
}

void scheduler(del3d_server_s* sv) {
    (void)hipSetDevice(sv->dev);
    std::unique_lock<std::mutex> lock(sv->mu);
    for (;;) {
        Slot* free_slot = nullptr;
        sv->cv_sched.wait(lock, [&] {
            if (sv->stop) return true;
            if (sv->queue.empty() || sv->blocked) return false;
            for (auto& s : sv->slots)
                if (!s->busy) {
                    free_slot = s.get();
                    return true;
                }
            return false;
        });
        if (sv->stop) break;
        launch_batch(sv, *free_slot, lock);
    }
}

// Queues a job (caller validated its buffers).
int enqueue_job(del3d_server_s* sv, const std::shared_ptr<Job>& j, int64_t* job_id) {
    std::lock_guard<std::mutex> lock(sv->mu);
    auto it = sv->sets.find(j->set);
    if (it == sv->sets.end()) return set_error(DEL3D_ERR_INVALID, "unknown set %lld", (long long)j->set);
    if (j->kind == kJobSort) j->n_q = it->second.n;
    ++it->second.refs;
    j->id = sv->next_job++;
    sv->jobs[j->id] = j;
    sv->queue.push_back(j);
    *job_id = j->id;
    sv->cv_sched.notify_one();
    return DEL3D_OK;
}

}  // namespace
}  // namespace d3

extern "C" {

int del3d_server_create(int64_t hbm_budget_bytes, int32_t n_slots, int64_t max_batch_queries,
                        del3d_server* out) {
    return guarded([&] {
        if (!out) return set_error(DEL3D_ERR_INVALID, "null out");
        *out = nullptr;
        if (hbm_budget_bytes <= 0 || n_slots < 1 || n_slots > 8 || max_batch_queries < 1 ||
            max_batch_queries > (int64_t)kChunk)
            return set_error(DEL3D_ERR_INVALID, "budget > 0, 1..8 slots, 1..2^26 queries per batch");
        auto sv = std::make_unique<del3d_server_s>();
        D3_HIP(hipGetDevice(&sv->dev));
        sv->budget = hbm_budget_bytes;
        sv->max_batch_q = max_batch_queries;
        for (int i = 0; i < n_slots; ++i) {
            auto sl = std::make_unique<Slot>();
            D3_HIP(hipStreamCreateWithFlags(&sl->s, hipStreamNonBlocking));
            D3_HIP(hipEventCreateWithFlags(&sl->done, hipEventDisableTiming));
            sv->slots.push_back(std::move(sl));
        }
        del3d_server_s* raw = sv.release();
        for (auto& sl : raw->slots) sl->worker = std::thread(slot_worker, raw, sl.get());
        raw->sched = std::thread(scheduler, raw);
        *out = raw;
        return DEL3D_OK;
    });
}

int del3d_server_put_points(del3d_server sv, const double* xyz, int64_t n_pts, int64_t* set_id) {
    return guarded([&] {
        if (!sv || !set_id || (n_pts > 0 && !xyz)) return set_error(DEL3D_ERR_INVALID, "null argument");
        if (n_pts < 1 || n_pts > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "bad point count");
        int64_t id = 0;
        double* dev = nullptr;
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            id = sv->next_set++;
            ResidentSet& r = sv->sets[id];
            r.host = xyz;
            r.n = n_pts;
            r.last_use = ++sv->tick;
            const int rc = guarded([&] {
                allocate_set(sv, r);
                return DEL3D_OK;
            });
            if (rc != DEL3D_OK) {
                sv->sets.erase(id);
                return rc;
            }
            ++r.pinned;
            dev = r.dev;
        }
        // the copy runs without the lock: submit, wait and the slots go on
        const int rc = guarded([&] {
            upload_set(xyz, dev, n_pts, thread_stream());
            return DEL3D_OK;
        });
        std::lock_guard<std::mutex> lock(sv->mu);
        ResidentSet& r = sv->sets[id];
        r.uploading = false;
        --r.pinned;
        sv->blocked = false;
        sv->cv_sched.notify_all();
        if (rc != DEL3D_OK) {
            (void)hipFree(r.dev);
            sv->used -= r.n * 24;
            sv->sets.erase(id);
            return rc;
        }
        *set_id = id;
        return DEL3D_OK;
    });
}

int del3d_server_drop_points(del3d_server sv, int64_t set_id) {
    return guarded([&] {
        if (!sv) return set_error(DEL3D_ERR_INVALID, "null server");
        std::lock_guard<std::mutex> lock(sv->mu);
        auto it = sv->sets.find(set_id);
        if (it == sv->sets.end()) return set_error(DEL3D_ERR_INVALID, "unknown set %lld", (long long)set_id);
        if (it->second.refs > 0 || it->second.uploading)
            return set_error(DEL3D_ERR_INVALID, "set %lld has pending jobs", (long long)set_id);
        if (it->second.dev) {
            D3_HIP(hipFree(it->second.dev));
            sv->used -= it->second.n * 24;
        }
        sv->sets.erase(it);
        return DEL3D_OK;
    });
}

int del3d_server_submit(del3d_server sv, int64_t set_id, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* sign_out, int64_t* job_id) {
    return guarded([&] {
        if (!sv || !job_id) return set_error(DEL3D_ERR_INVALID, "null argument");
        if (n_q < 1 || n_q > sv->max_batch_q)
            return set_error(DEL3D_ERR_INVALID, "a job holds 1..%lld queries", (long long)sv->max_batch_q);
        if (!op || !idx || !sign_out) return set_error(DEL3D_ERR_INVALID, "null buffer");
        auto j = std::make_shared<Job>();
        j->kind = kJobPredicates;
        j->set = set_id;
        j->op = op;
        j->idx = idx;
        j->n_q = n_q;
        j->out = sign_out;
        j->t_submit = Clock::now();
        return enqueue_job(sv, j, job_id);
    });
}

int del3d_server_submit_sort(del3d_server sv, int64_t set_id, int32_t* order_out, int64_t* job_id) {
    return guarded([&] {
        if (!sv || !job_id || !order_out) return set_error(DEL3D_ERR_INVALID, "null argument");
        auto j = std::make_shared<Job>();
        j->kind = kJobSort;
        j->set = set_id;
        j->order = order_out;
        j->t_submit = Clock::now();
        return enqueue_job(sv, j, job_id);
    });
}

int del3d_server_wait(del3d_server sv, int64_t job_id, double* latency_ms) {
    return guarded([&] {
        if (!sv) return set_error(DEL3D_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lock(sv->mu);
        auto it = sv->jobs.find(job_id);
        if (it == sv->jobs.end()) return set_error(DEL3D_ERR_INVALID, "unknown job %lld", (long long)job_id);
        auto j = it->second;
        sv->cv_done.wait(lock, [&] { return j->state >= 2; });
        sv->jobs.erase(it);
        if (latency_ms)
            *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
        if (j->rc != DEL3D_OK) return set_error(j->rc, "job %lld: %s", (long long)job_id, j->err.c_str());
        return DEL3D_OK;
    });
}

int del3d_server_info(del3d_server sv, int64_t* info) {
    return guarded([&] {
        if (!sv || !info) return set_error(DEL3D_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(sv->mu);
        int64_t resident = 0, nres = 0;
        for (auto& kv : sv->sets)
            if (kv.second.dev && !kv.second.uploading) {
                resident += kv.second.n * 24;
                ++nres;
            }
        info[0] = sv->used;
        info[1] = resident;
        info[2] = nres;
        info[3] = sv->stat_evictions;
        info[4] = sv->stat_uploads;
        info[5] = sv->stat_batches;
        info[6] = sv->stat_done;
        info[7] = sv->budget;
        return DEL3D_OK;
    });
}

int del3d_server_destroy(del3d_server sv) {
    if (!sv) return DEL3D_OK;
    {
        std::unique_lock<std::mutex> lock(sv->mu);
        // let queued and running jobs finish
        sv->cv_done.wait(lock, [&] {
            if (!sv->queue.empty()) return false;
            for (auto& s : sv->slots)
                if (s->busy) return false;
            return true;
        });
        sv->stop = true;
        sv->cv_sched.notify_all();
    }
    sv->sched.join();
    {
        std::lock_guard<std::mutex> lock(sv->mu);
        for (auto& s : sv->slots) {
            s->quit = true;
            s->cv.notify_one();
        }
    }
    for (auto& s : sv->slots) {
        if (s->worker.joinable()) s->worker.join();
        slot_release(sv, *s);
        (void)guarded([&] {
            slot_release_cache(sv, *s);
            return DEL3D_OK;
        });
        (void)hipEventDestroy(s->done);
        (void)hipStreamDestroy(s->s);
    }
    for (auto& kv : sv->sets)
        if (kv.second.dev) (void)hipFree(kv.second.dev);
    delete sv;
    return DEL3D_OK;
}

}  // extern "C"
