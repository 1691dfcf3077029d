// server.hip - a serving loop for predicate jobs (continuous batching).
//
// Clients register point sets (uploaded once, kept resident in HBM) and
// submit jobs: a batch of queries on one resident set, with host buffers.
// A scheduler thread admits queued jobs into batches and hands each batch to
// a free slot; each slot has its own thread, stream and device arenas, so the
// copies of one batch overlap the kernels and copies of others, and jobs that
// arrive while batches run are admitted as soon as a slot frees up. A
// batch is one launch sequence of the batched predicate kernels over its jobs,
// each job a set of the batch (per-set point pointers), so every job's signs
// are exactly those of del3d_predicates on its set alone.
//
// Memory: resident sets plus slot arenas stay within the HBM budget. When an
// upload or a batch's buffers would exceed it, sets no running batch uses are
// evicted least recently used first, then idle slots' buffers are freed; an
// evicted set is uploaded again from the client's host array (which must stay
// valid while the set is registered) when a job needs it. If memory is still
// short while other batches run, the jobs wait at the head of the queue for
// one to finish; if nothing else runs, a job that still does not fit fails
// with DEL3D_ERR_NOMEM, without affecting other jobs.
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
namespace {

struct ResidentSet {
    const double* host = nullptr;
    int64_t n = 0;
    double* dev = nullptr;  // nullptr while evicted
    int refs = 0;           // queued or running jobs (the set cannot be dropped)
    int pinned = 0;         // running jobs (the set cannot be evicted)
    uint64_t last_use = 0;
};

struct Job {
    int64_t id = 0, set = 0, n_q = 0;
    const uint8_t* op = nullptr;
    const int32_t* idx = nullptr;
    int8_t* out = nullptr;
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
    // pinned host staging
    int8_t* h_out = nullptr;
    const double** h_ptr = nullptr;
    int64_t* h_np = nullptr;
    int64_t* h_qoff = nullptr;
    int64_t* h_first = nullptr;
    PredTask* h_tasks = nullptr;
    int64_t cap_q = 0, cap_jobs = 0, cap_tasks = 0;
    int64_t bytes = 0;  // device bytes held
    bool busy = false;
    int64_t q_run = 0, n_tasks = 0, n_lds = 0;  // the batch in flight
    std::vector<std::shared_ptr<Job>> jobs;
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
    bool blocked = false;  // the queue head waits for memory held by running batches
    int64_t stat_evictions = 0, stat_uploads = 0, stat_batches = 0, stat_done = 0;
    std::thread sched;
};

namespace d3 {
namespace {

// Frees every buffer of a slot (device and pinned) and returns its bytes.
void slot_release(del3d_server_s* sv, Slot& sl) {
    for (void* p : {(void*)sl.d_op, (void*)sl.d_idx, (void*)sl.d_out, (void*)sl.d_queue,
                    (void*)sl.d_wide, (void*)sl.d_ctr, (void*)sl.d_ptr})
        if (p) (void)hipFree(p);
    for (void* p : {(void*)sl.h_out, (void*)sl.h_ptr})
        if (p) (void)hipHostFree(p);
    sl.d_op = nullptr;
    sl.d_idx = nullptr;
    sl.d_out = nullptr;
    sl.d_queue = sl.d_wide = nullptr;
    sl.d_ctr = nullptr;
    sl.d_ptr = nullptr;
    sl.d_np = sl.d_qoff = nullptr;
    sl.d_tasks = nullptr;
    sl.h_out = nullptr;
    sl.h_ptr = nullptr;
    sl.h_np = sl.h_qoff = nullptr;
    sl.h_tasks = nullptr;
    sv->used -= sl.bytes;
    sl.bytes = 0;
    sl.cap_q = sl.cap_jobs = sl.cap_tasks = 0;
}

// set pointers [j], point counts [j], query offsets [j+1], first points [j+1],
// grid summaries [2j int32], tasks [t]
size_t table_bytes(int64_t j, int64_t t) {
    return (size_t)(8 * j + 8 * j + 8 * (j + 1) + 8 * (j + 1) + 8 * j + t * (int64_t)sizeof(PredTask));
}

// Device bytes of a slot sized for q queries, j jobs and t tasks.
int64_t slot_bytes(int64_t q, int64_t j, int64_t t) {
    return q * (1 + 20 + 1 + 4 + 4) + (int64_t)sizeof(PredCounters) + (int64_t)table_bytes(j, t);
}

// Evicts resident sets no running job uses, least recently used first, then
// frees the buffers of idle slots (a slot being filled reallocates its own),
// until `need` more bytes fit the budget. Returns whether they do. Caller
// holds sv->mu.
bool make_room(del3d_server_s* sv, int64_t need) {
    while (sv->used + need > sv->budget) {
        int64_t victim = -1;
        uint64_t oldest = ~0ull;
        for (auto& kv : sv->sets) {
            const ResidentSet& r = kv.second;
            if (r.dev && r.pinned == 0 && r.last_use < oldest) {
                oldest = r.last_use;
                victim = kv.first;
            }
        }
        if (victim < 0) {
            // then the buffers of idle slots
            Slot* idle = nullptr;
            for (auto& sl : sv->slots)
                if (!sl->busy && sl->bytes > 0) idle = sl.get();
            if (!idle) return false;
            slot_release(sv, *idle);
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

// Makes a registered set resident (uploading it on the slot's stream).
// Caller holds sv->mu. Throws HipError on failure.
void ensure_resident(del3d_server_s* sv, ResidentSet& r, hipStream_t s) {
    if (r.dev) return;
    const int64_t bytes = r.n * 24;
    if (!make_room(sv, bytes))
        throw HipError{set_error(DEL3D_ERR_NOMEM, "point set of %lld bytes does not fit the HBM budget",
                                 (long long)bytes)};
    D3_HIP(hipMalloc((void**)&r.dev, (size_t)std::max<int64_t>(bytes, 24)));
    sv->used += bytes;
    ++sv->stat_uploads;
    copy_h2d(r.dev, r.host, (size_t)bytes, s);
}

// Grows a slot's buffers to hold q queries, j jobs, t tasks.
void slot_reserve(del3d_server_s* sv, Slot& sl, int64_t q, int64_t j, int64_t t) {
    if (q <= sl.cap_q && j <= sl.cap_jobs && t <= sl.cap_tasks) return;
    q = std::max(q, sl.cap_q);
    j = std::max(j, sl.cap_jobs);
    t = std::max(t, sl.cap_tasks);
    const int64_t want = slot_bytes(q, j, t);
    slot_release(sv, sl);
    if (!make_room(sv, want))
        throw HipError{set_error(DEL3D_ERR_NOMEM, "batch buffers of %lld bytes do not fit the HBM budget",
                                 (long long)want)};
    sv->used += want;
    sl.bytes = want;
    D3_HIP(hipMalloc((void**)&sl.d_op, (size_t)q));
    D3_HIP(hipMalloc((void**)&sl.d_idx, (size_t)q * 20));
    D3_HIP(hipMalloc((void**)&sl.d_out, (size_t)q));
    D3_HIP(hipMalloc((void**)&sl.d_queue, (size_t)q * 4));
    D3_HIP(hipMalloc((void**)&sl.d_wide, (size_t)q * 4));
    D3_HIP(hipMalloc((void**)&sl.d_ctr, sizeof(PredCounters)));
    // the batch tables live in one block (one copy per batch):
    // set pointers [j], point counts [j], query offsets [j+1], tasks [t]
    const size_t tab = table_bytes(j, t);
    char* dt = nullptr;
    D3_HIP(hipMalloc((void**)&dt, tab));
    sl.d_ptr = reinterpret_cast<const double**>(dt);
    sl.d_np = reinterpret_cast<int64_t*>(dt + 8 * j);
    sl.d_qoff = reinterpret_cast<int64_t*>(dt + 16 * j);
    sl.d_first = reinterpret_cast<int64_t*>(dt + 16 * j + 8 * (j + 1));
    sl.d_grid = reinterpret_cast<int32_t*>(dt + 16 * j + 16 * (j + 1));
    sl.d_tasks = reinterpret_cast<PredTask*>(dt + 24 * j + 16 * (j + 1));
    D3_HIP(hipHostMalloc((void**)&sl.h_out, (size_t)q, hipHostMallocDefault));
    char* ht = nullptr;
    D3_HIP(hipHostMalloc((void**)&ht, tab, hipHostMallocDefault));
    sl.h_ptr = reinterpret_cast<const double**>(ht);
    sl.h_np = reinterpret_cast<int64_t*>(ht + 8 * j);
    sl.h_qoff = reinterpret_cast<int64_t*>(ht + 16 * j);
    sl.h_first = reinterpret_cast<int64_t*>(ht + 16 * j + 8 * (j + 1));
    sl.h_tasks = reinterpret_cast<PredTask*>(ht + 24 * j + 16 * (j + 1));
    sl.cap_q = q;
    sl.cap_jobs = j;
    sl.cap_tasks = t;
}

constexpr int64_t kTaskQ = 16 * kGlobalThreads;  // queries per task (both kernels)

int64_t tasks_for(int64_t n_q) { return (n_q + kTaskQ - 1) / kTaskQ; }

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

// Runs a slot's batch: copies the jobs' inputs from the clients' host arrays
// (the runtime stages pageable memory itself), enqueues the kernels and the
// copy of the signs into pinned memory, waits, hands each job its signs.
void slot_run(del3d_server_s* sv, Slot* sl) {
    const int64_t q_run = sl->q_run, t = sl->n_tasks, n_lds = sl->n_lds;
    const size_t nj = sl->jobs.size();
    int rc = guarded([&] {
        hipStream_t s = sl->s;
        int64_t off = 0;
        for (auto& j : sl->jobs) {
            D3_HIP(hipMemcpyAsync(sl->d_op + off, j->op, (size_t)j->n_q, hipMemcpyHostToDevice, s));
            D3_HIP(hipMemcpyAsync(sl->d_idx + 5 * off, j->idx, (size_t)j->n_q * 20,
                                  hipMemcpyHostToDevice, s));
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
        D3_HIP(hipMemcpyAsync(sl->h_out, sl->d_out, (size_t)q_run, hipMemcpyDeviceToHost, s));
        D3_HIP(hipEventRecord(sl->done, s));
        D3_HIP(hipEventSynchronize(sl->done));
        return DEL3D_OK;
    });
    std::string err = rc ? last_error_slot() : std::string();
    if (rc) (void)hipStreamSynchronize(sl->s);  // let whatever was enqueued drain
    std::lock_guard<std::mutex> lock(sv->mu);
    int64_t off = 0;
    for (auto& j : sl->jobs) {
        if (rc == DEL3D_OK) {
            std::memcpy(j->out, sl->h_out + off, (size_t)j->n_q);
            finish_job(sv, j, DEL3D_OK, nullptr);
        } else {
            finish_job(sv, j, rc, err.c_str());
        }
        off += j->n_q;
    }
    sl->jobs.clear();
    sl->busy = false;
    sv->blocked = false;  // memory may have been freed
    sv->cv_done.notify_all();
    sv->cv_sched.notify_all();
}

// Builds and launches one batch from the queue into a free slot. Caller
// holds the lock; returns false if there was nothing to launch.
bool launch_batch(del3d_server_s* sv, Slot& sl, std::unique_lock<std::mutex>& lock) {
    std::vector<std::shared_ptr<Job>> take;
    int64_t q = 0;
    while (!sv->queue.empty()) {
        auto& j = sv->queue.front();
        if (!take.empty() && q + j->n_q > sv->max_batch_q) break;
        q += j->n_q;
        take.push_back(j);
        sv->queue.pop_front();
    }
    if (take.empty()) return false;
    // Residency and buffers, evicting unused sets if needed. When memory is
    // short while other batches run, the jobs go back to the head of the
    // queue until one of those batches finishes; when nothing else runs, a
    // job that still does not fit fails alone.
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
    std::vector<std::shared_ptr<Job>> run;
    int64_t q_run = 0, t_run = 0;
    for (size_t k = 0; k < take.size(); ++k) {
        auto& j = take[k];
        ResidentSet& r = sv->sets[j->set];
        const int rc = guarded([&] {
            ensure_resident(sv, r, sl.s);
            return DEL3D_OK;
        });
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
    if (run.empty()) {
        sv->cv_done.notify_all();
        return true;
    }
    int rc = guarded([&] {
        slot_reserve(sv, sl, q_run, (int64_t)run.size(), t_run);
        return DEL3D_OK;
    });
    if (rc == DEL3D_ERR_NOMEM && other_slot_busy(sv, &sl)) {
        defer(run, 0);
        return true;
    }
    if (rc == DEL3D_ERR_NOMEM && run.size() > 1) {
        // nothing else runs: try the first job alone, the others wait
        defer(run, 1);
        sv->blocked = false;
        q_run = run[0]->n_q;
        t_run = tasks_for(q_run);
        rc = guarded([&] {
            slot_reserve(sv, sl, q_run, 1, t_run);
            return DEL3D_OK;
        });
    }
    if (rc != DEL3D_OK) {
        for (auto& j : run) finish_job(sv, j, rc, last_error_slot().c_str());
        sv->cv_done.notify_all();
        return true;
    }
    // host tables and inputs (the set pointers are stable while pinned)
    int64_t off = 0, t = 0;
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
    const int64_t n_lds = std::count_if(sl.h_tasks, sl.h_tasks + t, [](const PredTask& x) { return x.pad == 1; });
    for (int64_t i = 0; i < t; ++i) sl.h_tasks[i].pad = 0;
    sl.jobs = run;
    sl.busy = true;
    sl.q_run = q_run;
    sl.n_tasks = t;
    sl.n_lds = n_lds;
    ++sv->stat_batches;
    // the slot's own thread copies the inputs, runs and completes the batch,
    // so the batches of different slots move their data concurrently
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
        std::lock_guard<std::mutex> lock(sv->mu);
        const int64_t id = sv->next_set++;
        ResidentSet& r = sv->sets[id];
        r.host = xyz;
        r.n = n_pts;
        r.last_use = ++sv->tick;
        const int rc = guarded([&] {
            ensure_resident(sv, r, thread_stream());
            return DEL3D_OK;
        });
        if (rc != DEL3D_OK) {
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
        if (it->second.refs > 0) return set_error(DEL3D_ERR_INVALID, "set %lld has pending jobs", (long long)set_id);
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
        j->set = set_id;
        j->op = op;
        j->idx = idx;
        j->n_q = n_q;
        j->out = sign_out;
        j->t_submit = Clock::now();
        std::lock_guard<std::mutex> lock(sv->mu);
        auto it = sv->sets.find(set_id);
        if (it == sv->sets.end()) return set_error(DEL3D_ERR_INVALID, "unknown set %lld", (long long)set_id);
        ++it->second.refs;
        j->id = sv->next_job++;
        sv->jobs[j->id] = j;
        sv->queue.push_back(j);
        *job_id = j->id;
        sv->cv_sched.notify_one();
        return DEL3D_OK;
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
            if (kv.second.dev) {
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
        (void)hipEventDestroy(s->done);
        (void)hipStreamDestroy(s->s);
    }
    for (auto& kv : sv->sets)
        if (kv.second.dev) (void)hipFree(kv.second.dev);
    delete sv;
    return DEL3D_OK;
}

}  // extern "C"
