// Emulated (CPU, AddressSanitizer) run of the GPU spatial sort, checked
// against the oracle (test infrastructure). Usage: emu_sort N [N ...]
// Environment: EMU_DEFER=1 (queued kernels, immediate pageable copies and
// frees), EMU_BLOCK_ORDER=0|1|2.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../oracle/oracle.h"

extern "C" int emu_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order);
extern "C" int emu_predicates_stats(const double* xyz, int64_t n_pts, const uint8_t* op,
                                    const int32_t* idx, int64_t n_q, int8_t* out, int64_t* st);

// xorshift64*, for the predicate workloads
static uint64_t g_rng = 88172645463325252ull;
static uint64_t rnd() {
    g_rng ^= g_rng >> 12;
    g_rng ^= g_rng << 25;
    g_rng ^= g_rng >> 27;
    return g_rng * 2685821657736338717ull;
}

// Mixed queries over a point set, with aliased slots; returns the failures.
static int check_predicates(const char* name, const std::vector<double>& xyz, int64_t nq) {
    const int64_t n = (int64_t)xyz.size() / 3;
    std::vector<uint8_t> op(nq);
    std::vector<int32_t> idx(5 * nq);
    for (int64_t q = 0; q < nq; ++q) {
        op[q] = (uint8_t)(rnd() % 3);
        for (int k = 0; k < 5; ++k) idx[5 * q + k] = (int32_t)(rnd() % n);
        if (q % 7 == 0) idx[5 * q + 4] = idx[5 * q];
    }
    std::vector<int8_t> got(nq), want(nq);
    int64_t st[6] = {0, 0, 0, 0, 0, 0};
    if (emu_predicates_stats(xyz.data(), n, op.data(), idx.data(), nq, got.data(), st) != 0) {
        std::printf("%s error\n", name);
        return 1;
    }
    orc_int_pred_batch(xyz.data(), n, op.data(), idx.data(), nq, want.data());
    const bool ok = got == want && st[0] + st[1] + st[2] + st[3] == nq;
    std::printf("pred %s %s (filter %lld cert %lld regs %lld wide %lld perturbed %lld)\n", name,
                ok ? "ok" : "MISMATCH", (long long)st[0], (long long)st[1], (long long)st[2],
                (long long)st[3], (long long)st[4]);
    return ok ? 0 : 1;
}
extern "C" int emu_predicates_forced(const double* xyz, int64_t n_pts, const uint8_t* op,
                                     const int32_t* idx, int64_t n_q, int8_t* out, int path);
extern "C" int emu_predicates_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                                   const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                                   int8_t* out);

extern "C" int emu_plan_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                             const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                             int8_t* out, int glob_q, int binned);
extern "C" int emu_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op,
                                      const int32_t* idx, int64_t n_q, int8_t* out);
extern "C" int emu_predicates_small(const double* xyz, int64_t n_pts, const uint8_t* op,
                                    const int32_t* idx, int64_t n_q, int8_t* out);

// The one-pass small-batch kernel against the integer oracle.
static int check_small(const char* name, const std::vector<double>& xyz, int64_t nq) {
    const int64_t n = (int64_t)xyz.size() / 3;
    std::vector<uint8_t> op(nq);
    std::vector<int32_t> idx(5 * nq);
    for (int64_t q = 0; q < nq; ++q) {
        op[q] = (uint8_t)(rnd() % 3);
        for (int k = 0; k < 5; ++k) idx[5 * q + k] = (int32_t)(rnd() % n);
        if (q % 7 == 0) idx[5 * q + 4] = idx[5 * q];
        if (q % 97 == 0) idx[5 * q + 2] = (int32_t)n;  // an index outside the set: sign 0
        if (q % 89 == 0) op[q] = 9;                     // an unknown op: sign 0
    }
    std::vector<int8_t> got(nq, 7), want(nq);
    const int rc = emu_predicates_small(xyz.data(), n, op.data(), idx.data(), nq, got.data());
    for (int64_t q = 0; q < nq; ++q) {
        bool bad = op[q] > 2;
        for (int k = 0; k < (op[q] == 0 ? 4 : 5); ++k) bad = bad || idx[5 * q + k] >= n;
        if (bad) {
            op[q] = 0;  // the oracle's sign is replaced below
            for (int k = 0; k < 5; ++k) idx[5 * q + k] = 0;
        }
    }
    orc_int_pred_batch(xyz.data(), n, op.data(), idx.data(), nq, want.data());
    bool ok = rc == 0;
    for (int64_t q = 0; q < nq && ok; ++q) {
        const bool bad = q % 97 == 0 || q % 89 == 0;
        ok = bad ? got[q] == 0 : got[q] == want[q];
    }
    std::printf("small %s %s\n", name, ok ? "ok" : "MISMATCH");
    return ok ? 0 : 1;
}

// The allocation-free form against the integer oracle, on consecutive calls
// sharing its fixed scratch (the second on a smaller set).
static int check_noalloc(const char* name, const std::vector<double>& xyz, int64_t nq) {
    const int64_t n = (int64_t)xyz.size() / 3;
    std::vector<uint8_t> op(nq);
    std::vector<int32_t> idx(5 * nq);
    for (int64_t q = 0; q < nq; ++q) {
        op[q] = (uint8_t)(rnd() % 3);
        for (int k = 0; k < 5; ++k) idx[5 * q + k] = (int32_t)(rnd() % n);
        if (q % 7 == 0) idx[5 * q + 4] = idx[5 * q];
    }
    std::vector<int8_t> got(nq, 9), want(nq);
    const int rc = emu_predicates_noalloc(xyz.data(), n, op.data(), idx.data(), nq, got.data());
    orc_int_pred_batch(xyz.data(), n, op.data(), idx.data(), nq, want.data());
    const bool ok = rc == 0 && got == want;
    std::printf("noalloc %s %s\n", name, ok ? "ok" : "MISMATCH");
    return ok ? 0 : 1;
}

// The same queries through the LDS kernel (path 1) and the global kernel
// (path 2), each against the integer oracle.
static int check_paths(const char* name, const std::vector<double>& xyz, int64_t nq) {
    const int64_t n = (int64_t)xyz.size() / 3;
    std::vector<uint8_t> op(nq);
    std::vector<int32_t> idx(5 * nq);
    for (int64_t q = 0; q < nq; ++q) {
        op[q] = (uint8_t)(rnd() % 3);
        for (int k = 0; k < 5; ++k) idx[5 * q + k] = (int32_t)(rnd() % n);
        if (q % 5 == 0) idx[5 * q + 4] = idx[5 * q + 1];
    }
    idx[5 * 3 + 2] = (int32_t)n;  // one out-of-range index: sign 0, no read outside the set
    std::vector<int8_t> want(nq);
    orc_int_pred_batch(xyz.data(), n, op.data(), idx.data(), nq, want.data());
    want[3] = 0;
    int fails = 0;
    for (int path = 1; path <= 2; ++path) {
        std::vector<int8_t> got(nq, 9);
        const int rc = emu_predicates_forced(xyz.data(), n, op.data(), idx.data(), nq, got.data(), path);
        const bool ok = rc == 0 && got == want;
        std::printf("pred %s path %d %s\n", name, path, ok ? "ok" : "MISMATCH");
        fails += ok ? 0 : 1;
    }
    return fails;
}

// Several sets of different sizes (LDS-sized and larger) in one call.
static int check_sets() {
    const int64_t sizes[5] = {7, 300, 5632, 5633, 9000};
    const int64_t nqs[5] = {40, 1500, 3000, 2500, 2100};
    std::vector<double> xyz;
    std::vector<int64_t> po = {0}, qo = {0};
    std::vector<uint8_t> op;
    std::vector<int32_t> idx;
    std::vector<int8_t> want;
    for (int s = 0; s < 5; ++s) {
        std::vector<double> p(3 * (size_t)sizes[s]);
        for (auto& c : p) c = (double)(int)(rnd() % 9) * 0.25;  // coarse grid: many degeneracies
        std::vector<uint8_t> o(nqs[s]);
        std::vector<int32_t> ix(5 * nqs[s]);
        for (int64_t q = 0; q < nqs[s]; ++q) {
            o[q] = (uint8_t)(rnd() % 3);
            for (int k = 0; k < 5; ++k) ix[5 * q + k] = (int32_t)(rnd() % sizes[s]);
        }
        std::vector<int8_t> w(nqs[s]);
        orc_int_pred_batch(p.data(), sizes[s], o.data(), ix.data(), nqs[s], w.data());
        xyz.insert(xyz.end(), p.begin(), p.end());
        op.insert(op.end(), o.begin(), o.end());
        idx.insert(idx.end(), ix.begin(), ix.end());
        want.insert(want.end(), w.begin(), w.end());
        po.push_back(po.back() + sizes[s]);
        qo.push_back(qo.back() + nqs[s]);
    }
    std::vector<int8_t> got(want.size(), 9);
    const int rc = emu_predicates_sets(xyz.data(), po.data(), 5, op.data(), idx.data(), qo.data(), got.data());
    bool ok = rc == 0 && got == want;
    std::printf("pred sets %s\n", ok ? "ok" : "MISMATCH");
    // every set through the global kernel in tasks of 256 queries, with and
    // without XCD bins (5 sets in 8 bins: most workgroups start on an empty bin)
    for (int binned = 0; binned < 2; ++binned) {
        std::vector<int8_t> g2(want.size(), 9);
        const int rc2 = emu_plan_sets(xyz.data(), po.data(), 5, op.data(), idx.data(), qo.data(), g2.data(),
                                      256, binned);
        const bool ok2 = rc2 == 0 && g2 == want;
        std::printf("pred sets plan binned=%d %s\n", binned, ok2 ? "ok" : "MISMATCH");
        ok = ok && ok2;
    }
    return ok ? 0 : 1;
}
// The serving loop: jobs from two client threads over four sets whose sizes
// exceed the HBM budget together (so idle sets are evicted and uploaded
// again), each job's signs against the integer oracle; a set larger than the
// budget is refused.
#include <thread>
#include "../../include/del3d_amd.h"
static int check_server() {
    const int64_t sizes[4] = {50, 300, 6000, 3000};
    std::vector<std::vector<double>> pts(4);
    for (int s = 0; s < 4; ++s) {
        pts[s].resize(3 * (size_t)sizes[s]);
        for (auto& c : pts[s]) c = (double)(int)(rnd() % 7) * 0.5;
    }
    del3d_server sv = nullptr;
    if (del3d_server_create(430000, 2, 4096, &sv) != 0) { std::printf("server create failed\n"); return 1; }
    int64_t ids[4];
    for (int s = 0; s < 4; ++s)
        if (del3d_server_put_points(sv, pts[s].data(), sizes[s], &ids[s]) != 0) {
            std::printf("put %d failed: %s\n", s, del3d_last_error());
            return 1;
        }
    std::vector<double> huge(3 * 20000, 1.0);
    int64_t hid = 0;
    int fails = 0;
    if (del3d_server_put_points(sv, huge.data(), 20000, &hid) != DEL3D_ERR_NOMEM) {
        std::printf("oversized set admitted\n");
        ++fails;
    }
    const int kJobs = 24;
    struct J { int set; std::vector<uint8_t> op; std::vector<int32_t> idx; std::vector<int8_t> out, want; int64_t id; };
    std::vector<J> jobs(kJobs);
    for (int k = 0; k < kJobs; ++k) {
        J& j = jobs[k];
        j.set = (int)(rnd() % 4);
        const int64_t nq = 1 + (int64_t)(rnd() % 3000);
        j.op.resize(nq);
        j.idx.resize(5 * nq);
        for (int64_t q = 0; q < nq; ++q) {
            j.op[q] = (uint8_t)(rnd() % 3);
            for (int c = 0; c < 5; ++c) j.idx[5 * q + c] = (int32_t)(rnd() % sizes[j.set]);
        }
        j.out.assign(nq, 9);
        j.want.resize(nq);
        orc_int_pred_batch(pts[j.set].data(), sizes[j.set], j.op.data(), j.idx.data(), nq, j.want.data());
    }
    auto client = [&](int first) {
        for (int k = first; k < kJobs; k += 2) {
            J& j = jobs[k];
            if (del3d_server_submit(sv, ids[j.set], j.op.data(), j.idx.data(), (int64_t)j.op.size(),
                                    j.out.data(), &j.id) != 0) j.id = -1;
        }
        for (int k = first; k < kJobs; k += 2) {
            double ms = 0;
            if (jobs[k].id < 0 || del3d_server_wait(sv, jobs[k].id, &ms) != 0) {
                std::printf("server job %d failed: %s\n", k, del3d_last_error());
                jobs[k].out.assign(1, 9);
            }
        }
    };
    // sort jobs beside them, on the two smallest sets (a sort's scratch takes
    // most of the budget, so they wait for memory and evict)
    std::vector<int32_t> sort_out[3], sort_want[3];
    const int sort_set[3] = {0, 1, 0};
    int64_t sort_id[3];
    for (int k = 0; k < 3; ++k) {
        const int64_t n = sizes[sort_set[k]];
        sort_out[k].assign(n, -1);
        sort_want[k].resize(n);
        for (int64_t i = 0; i < n; ++i) sort_want[k][i] = (int32_t)i;
        orc_spatial_sort(pts[sort_set[k]].data(), n, sort_want[k].data());
        if (del3d_server_submit_sort(sv, ids[sort_set[k]], sort_out[k].data(), &sort_id[k]) != 0) {
            std::printf("sort submit failed: %s\n", del3d_last_error());
            ++fails;
            sort_id[k] = -1;
        }
    }
    std::thread t0(client, 0), t1(client, 1);
    t0.join();
    t1.join();
    for (int k = 0; k < 3; ++k) {
        double ms = 0;
        if (sort_id[k] < 0 || del3d_server_wait(sv, sort_id[k], &ms) != 0 || sort_out[k] != sort_want[k]) {
            std::printf("server sort job %d MISMATCH (%s)\n", k, del3d_last_error());
            ++fails;
        }
    }
    for (int k = 0; k < kJobs; ++k)
        if (jobs[k].out != jobs[k].want) {
            std::printf("server job %d (set %d) MISMATCH\n", k, jobs[k].set);
            ++fails;
        }
    int64_t info[8];
    del3d_server_info(sv, info);
    std::printf("server jobs %d batches %lld evictions %lld uploads %lld used %lld %s\n", kJobs,
                (long long)info[5], (long long)info[3], (long long)info[4], (long long)info[0],
                fails ? "MISMATCH" : "ok");
    if (info[3] == 0) {
        std::printf("server: no eviction happened\n");
        ++fails;
    }
    del3d_server_destroy(sv);
    return fails;
}
extern "C" int emu_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                                     int32_t* order);

extern "C" int emu_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op,
                               const int32_t* idx, int64_t n_q, int8_t* out);

// The coplanar batch (fast kernel, queue, exact kernel) on a cocircular ring
// mapped onto a tilted plane, an integer grid and random points, all ops, with
// aliased slots, plus queries with bad indices or op codes (result 0).
static int check_coplanar() {
    static const int ring[12][2] = {{5, 0}, {4, 3}, {3, 4}, {0, 5}, {-3, 4}, {-4, 3},
                                    {-5, 0}, {-4, -3}, {-3, -4}, {0, -5}, {3, -4}, {4, -3}};
    std::vector<double> xyz;
    for (auto& r : ring) {  // (x, y, 0) mapped by [[1,2,2],[2,1,-2],[2,-2,1]]
        const double x = r[0], y = r[1];
        xyz.insert(xyz.end(), {x + 2 * y, 2 * x + y, 2 * x - 2 * y});
    }
    for (int i = -3; i <= 3; ++i)
        for (int j = -3; j <= 3; ++j) xyz.insert(xyz.end(), {(double)i, (double)j, (double)(i + j)});
    for (int i = 0; i < 40; ++i)
        for (int k = 0; k < 3; ++k) xyz.push_back((double)(rnd() >> 11) * 0x1p-53);
    // a wide exponent span (more limbs than the lane-parallel exact pass takes)
    for (int i = 0; i < 20; ++i)
        for (int k = 0; k < 3; ++k)
            xyz.push_back(std::ldexp((double)(rnd() >> 11) * 0x1p-53, (int)(rnd() % 200) - 100));
    const int64_t n = (int64_t)xyz.size() / 3, nq = 30000;
    std::vector<uint8_t> op(nq);
    std::vector<int32_t> idx(5 * nq, 0);
    for (int64_t q = 0; q < nq; ++q) {
        op[q] = (uint8_t)(rnd() % 9);
        // mostly within one group (ring, grid or random) so the queries meet
        // the predicates' preconditions; some across groups
        const int64_t g = rnd() % 5, lo = g == 0 ? 0 : g == 1 ? 12 : g == 2 ? 61 : g == 3 ? 101 : 0;
        const int64_t w = g == 0 ? 12 : g == 1 ? 49 : g == 2 ? 40 : g == 3 ? 20 : n;
        for (int k = 0; k < 4; ++k) idx[5 * q + k] = (int32_t)(lo + rnd() % w);
        if (q % 9 == 0) idx[5 * q + 3] = idx[5 * q + (rnd() % 3)];
    }
    std::vector<int8_t> got(nq), want(nq);
    if (emu_predicates2(xyz.data(), n, op.data(), idx.data(), nq, got.data()) != 0) {
        std::printf("coplanar error\n");
        return 1;
    }
    orc_pred2_batch(xyz.data(), n, op.data(), idx.data(), nq, want.data());
    // bad inputs: an index out of range, a negative index, unknown op codes
    const uint8_t bop[4] = {5, 0, 8, 200};
    const int32_t bidx[20] = {0, 1, 2, (int32_t)n, 0, -1, 1, 2, 0, 0, 0, 1, 2, 3, 0, 0, 1, 2, 3, 0};
    int8_t bout[4] = {9, 9, 9, 9};
    const int rc = emu_predicates2(xyz.data(), n, bop, bidx, 4, bout);
    const bool ok = got == want && rc == 0 && bout[0] == 0 && bout[1] == 0 && bout[2] == 0 &&
                    bout[3] == 0;
    int64_t st[3];
    orc_last_stage_counts(st);
    std::printf("coplanar %s (oracle: filter %lld exact %lld perturbed %lld)\n", ok ? "ok" : "MISMATCH",
                (long long)st[0], (long long)st[1], (long long)st[2]);
    return ok ? 0 : 1;
}

extern "C" int emu_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n,
                                 int lazy, double* out);

// Circumcentres, lazy and exact, on random cells and slivers (the lazy pass
// refuses those and queues them for the exact kernel), plus a bad index.
static int check_circumcenters() {
    std::vector<double> xyz;
    for (int i = 0; i < 200; ++i)
        for (int k = 0; k < 3; ++k) xyz.push_back((double)(rnd() >> 11) * 0x1p-53);
    for (int i = 0; i < 50; ++i) {  // near-coplanar quadruples: z = x/4 + y/2 + tiny
        for (int j = 0; j < 4; ++j) {
            const double x = (double)(rnd() >> 11) * 0x1p-53, y = (double)(rnd() >> 11) * 0x1p-53;
            xyz.insert(xyz.end(), {x, y, 0.25 * x + 0.5 * y + (j == 3 ? 1e-13 : 0.0)});
        }
    }
    for (int i = 0; i < 100; ++i)  // mixed scales: expansions outgrow the small workspace
        for (int k = 0; k < 3; ++k)
            xyz.push_back(std::ldexp((double)(rnd() >> 11) * 0x1p-53 + 0.5, (int)(rnd() % 120) - 60));
    const int64_t n = (int64_t)xyz.size() / 3;
    std::vector<int32_t> cells;
    for (int c = 0; c < 600; ++c)
        for (int k = 0; k < 4; ++k) cells.push_back((int32_t)(rnd() % 200));
    for (int c = 0; c < 300; ++c)
        for (int k = 0; k < 4; ++k) cells.push_back((int32_t)(400 + rnd() % 100));
    for (int c = 0; c < 50; ++c)
        for (int k = 0; k < 4; ++k) cells.push_back(200 + 4 * c + k);
    cells.insert(cells.end(), {0, 1, 2, (int32_t)n});  // out of range: NaN
    const int64_t nc = (int64_t)cells.size() / 4;
    int fails = 0;
    for (int lazy = 0; lazy < 2; ++lazy) {
        std::vector<double> got(3 * nc), want(3 * nc);
        if (emu_circumcenters(xyz.data(), n, cells.data(), nc, lazy, got.data()) != 0) return 1;
        orc_circumcenters(xyz.data(), cells.data(), nc - 1, lazy, want.data());
        bool ok = std::memcmp(got.data(), want.data(), sizeof(double) * 3 * (nc - 1)) == 0;
        for (int k = 0; k < 3; ++k) ok = ok && std::isnan(got[3 * (nc - 1) + k]);
        std::printf("circumcenters %s %s\n", lazy ? "lazy" : "exact", ok ? "ok" : "MISMATCH");
        fails += ok ? 0 : 1;
    }
    return fails;
}

// One sort of n points (order = 0..n-1, or the given order) against the oracle.
static int check_sort(const char* name, const std::vector<double>& xyz, int64_t n,
                      const std::vector<int32_t>* order = nullptr) {
    std::vector<int32_t> got, want;
    if (order) {
        got = want = *order;
    } else {
        got.resize(n);
        for (int64_t i = 0; i < n; ++i) got[i] = (int32_t)i;
        want = got;
    }
    const int64_t m = (int64_t)got.size();
    if (emu_spatial_sort(xyz.data(), n, got.data(), m) != 0) {
        std::printf("sort %s error: %s\n", name, del3d_last_error());
        return 1;
    }
    orc_spatial_sort(xyz.data(), m, want.data());
    const bool ok = got == want;
    std::printf("sort %s %s\n", name, ok ? "ok" : "MISMATCH");
    return ok ? 0 : 1;
}

// Tie-heavy inputs, an order that is a permuted subset, and a bad index.
static int check_sort_cases() {
    int fails = 0;
    {
        std::vector<double> same;
        for (int i = 0; i < 4700; ++i) same.insert(same.end(), {0.25, 0.5, 0.75});
        fails += check_sort("identical 4700", same, 4700);
    }
    {
        std::vector<double> lat;
        for (int i = 0; i < 20; ++i)
            for (int j = 0; j < 20; ++j)
                for (int k = 0; k < 20; ++k) lat.insert(lat.end(), {(double)i, (double)j, (double)k});
        fails += check_sort("lattice 20^3", lat, 8000);
    }
    {
        std::vector<double> few;
        for (int i = 0; i < 12000; ++i)
            for (int k = 0; k < 3; ++k) few.push_back((double)(rnd() % 3));
        fails += check_sort("3 values per axis n=12000", few, 12000);
    }
    {
        std::vector<double> pts(3 * 9000 + 3);
        orc_pts_random_cube(9000, 3u, pts.data());
        std::vector<int32_t> ord;
        for (int i = 0; i < 9000; ++i)
            if (rnd() % 3) ord.push_back(i);
        for (size_t i = ord.size(); i > 1; --i) std::swap(ord[i - 1], ord[rnd() % i]);
        fails += check_sort("subset order", pts, 9000, &ord);
        std::vector<int32_t> bad = ord;
        bad[17] = 9000;
        const int rc = emu_spatial_sort(pts.data(), 9000, bad.data(), (int64_t)bad.size());
        const bool ok = rc == DEL3D_ERR_INVALID;
        std::printf("sort bad index %s (rc %d)\n", ok ? "ok" : "MISMATCH", rc);
        fails += ok ? 0 : 1;
    }
    return fails;
}

// Query counts of the predicate checks, divided by EMU_QUERY_DIV (the
// ThreadSanitizer run, where every emulated lane is an OS thread, uses 3:
// still several tiles per workgroup).
static int nq(int n) {
    const char* d = std::getenv("EMU_QUERY_DIV");
    const int k = d ? std::atoi(d) : 1;
    return k > 1 ? n / k : n;
}

int main(int argc, char** argv) {
    int fails = 0;
    // the serving loop alone (the ThreadSanitizer run)
    if (std::getenv("EMU_SERVER_ONLY")) return check_server() ? 1 : 0;
    if (!std::getenv("EMU_SORT_ONLY")) {
        std::vector<double> lat;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j)
                for (int k = 0; k < 5; ++k) lat.insert(lat.end(), {(double)i, (double)j, (double)k});
        fails += check_predicates("lattice", lat, nq(3000));
        std::vector<double> wide;
        for (int i = 0; i < 60; ++i)
            for (int k = 0; k < 3; ++k) {
                const double m = (double)(int64_t)(rnd() % 2000001) - 1000000.0;
                wide.push_back(std::ldexp(m + 0.5, (int)(rnd() % 400) - 200));
            }
        fails += check_predicates("wide", wide, nq(1500));
        // spans of 2^60 .. 2^119: 3-4 limbs per difference (the register
        // kernel for the wide queue), with near-degenerate queries from
        // near-coplanar and near-cospherical points
        std::vector<double> mid;
        for (int i = 0; i < 80; ++i) {
            const double u = (double)(rnd() % 1000003) / 1000003.0, v = (double)(rnd() % 999983) / 999983.0;
            const double s = std::ldexp(1.0, (int)(rnd() % 60) - 40);
            if (i % 2) mid.insert(mid.end(), {u * s, v * s, (u + v) * s});            // on a plane
            else mid.insert(mid.end(), {std::cos(6.3 * u) * s, std::sin(6.3 * u) * s, v * 1e-3});
        }
        fails += check_predicates("mid", mid, nq(3000));
        std::vector<double> q48;
        for (int i = 0; i < 200; ++i)
            for (int k = 0; k < 3; ++k) q48.push_back((double)(rnd() >> 16) * 3.552713678800501e-15);
        fails += check_predicates("grid48", q48, nq(2000));
        // above the LDS threshold, packable: the packed LDS kernel
        std::vector<double> lat20;
        for (int i = 0; i < 20; ++i)
            for (int j = 0; j < 20; ++j)
                for (int k = 0; k < 20; ++k) lat20.insert(lat20.end(), {i * 0.5, j * 0.25, k - 10.0});
        fails += check_predicates("lattice20 (packed)", lat20, nq(3000));
        // above the packed kernel's capacity, packable: the global kernel's
        // packed variant; moved by 1000, exact in float but too wide to pack:
        // the global kernel's float copy
        std::vector<double> lat34, lat34f;
        for (int i = 0; i < 34; ++i)
            for (int j = 0; j < 34; ++j)
                for (int k = 0; k < 34; ++k) {
                    lat34.insert(lat34.end(), {i * 0.5, j * 0.25, k - 10.0});
                    lat34f.insert(lat34f.end(), {i * 0.5 + 1000.0, j * 0.25, k - 10.0});
                }
        fails += check_predicates("lattice34 (global packed)", lat34, nq(3000));
        fails += check_predicates("lattice34+1000 (float copy)", lat34f, nq(3000));
        fails += check_small("lattice", lat, nq(3000));
        fails += check_small("wide", wide, nq(1500));
        fails += check_small("mid", mid, nq(3000));
        fails += check_small("grid48", q48, nq(2000));
        fails += check_noalloc("lattice34 (global packed)", lat34, nq(4000));
        fails += check_noalloc("lattice20 (packed)", lat20, nq(3000));
        fails += check_noalloc("grid48 (LDS)", q48, nq(2000));
        fails += check_paths("lattice", lat, nq(5000));
        fails += check_paths("wide", wide, nq(2500));
        fails += check_sets();
        fails += check_coplanar();
        fails += check_circumcenters();
        if (!std::getenv("EMU_DEFER") || std::atoi(std::getenv("EMU_DEFER")) == 0) fails += check_server();
        if (std::getenv("EMU_PRED_ONLY")) return fails ? 1 : 0;
    }
    for (int a = 1; a < argc; ++a) {
        const int64_t n = std::atoll(argv[a]);
        std::vector<double> xyz(3 * (size_t)n + 3);
        orc_pts_random_cube(n, 7u, xyz.data());
        // every fourth point snapped to a coarse grid: ties on every axis
        for (int64_t i = 0; i < n; i += 4)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = (double)(int)(xyz[3 * i + k] * 4) / 4;
        fails += check_sort(("random+grid n=" + std::to_string(n)).c_str(), xyz, n);
    }
    if (std::getenv("EMU_SORT_CASES")) fails += check_sort_cases();
    if (std::getenv("EMU_SKIP_SETS")) return fails ? 1 : 0;
    // sets in one call, on both sides of the LDS and subtree thresholds
    const int64_t sizes[8] = {1500, 3, 700, 0, 9, 5000, 1, 9100};
    std::vector<double> xyz;
    std::vector<int64_t> off = {0};
    std::vector<int32_t> got, want;
    for (int s = 0; s < 8; ++s) {
        std::vector<double> p(3 * (size_t)sizes[s] + 3);
        orc_pts_random_cube(sizes[s], 30u + s, p.data());
        if (s == 7)  // a lattice: masses of equal keys in the device rounds
            for (int64_t i = 0; i < sizes[s]; ++i)
                for (int k = 0; k < 3; ++k) p[3 * i + k] = (double)((i >> (4 * k)) & 15);
        xyz.insert(xyz.end(), p.begin(), p.begin() + 3 * sizes[s]);
        std::vector<int32_t> o(sizes[s]);
        for (int64_t i = 0; i < sizes[s]; ++i) o[i] = (int32_t)i;
        got.insert(got.end(), o.begin(), o.end());
        orc_spatial_sort(p.data(), sizes[s], o.data());
        want.insert(want.end(), o.begin(), o.end());
        off.push_back(off.back() + sizes[s]);
    }
    const bool sets_ok = emu_spatial_sort_sets(xyz.data(), off.data(), 8, got.data()) == 0 && got == want;
    std::printf("sets %s\n", sets_ok ? "ok" : "MISMATCH");
    fails += sets_ok ? 0 : 1;
    return fails ? 1 : 0;
}
