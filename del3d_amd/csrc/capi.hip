// capi.hip - the C ABI of include/del3d_amd.h: argument checks, error
// reporting, host-memory and scalar forms over the batched kernels.
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "host_scalar.h"
#include "predicates.h"

namespace d3 {

void spatial_sort_device(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                         hipStream_t s);
void spatial_sort_sets_device(const double* xyz, const int64_t* off, int64_t n_sets,
                              int32_t* order, hipStream_t s);

std::string& last_error_slot() {
    thread_local std::string msg;
    return msg;
}

int device_cu_count() {
    static std::mutex mu;
    static std::vector<int> cache;
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    if ((int)cache.size() <= dev) cache.resize(dev + 1, 0);
    if (cache[dev] == 0) {
        int cu = 0;
        D3_HIP(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = cu > 0 ? cu : 1;
    }
    return cache[dev];
}

// ---- scratch cache, pinned staging, fill -----------------------------------
// Device scratch blocks are cached per (device, stream) after a call, so the
// next call on that stream reuses them without hipMalloc. The cache holds at
// most kCacheCapPerStream bytes per stream (a block given back beyond that is
// freed), is emptied and retried once when hipMalloc runs out of memory, and
// del3d_release_cache() empties it on request.
namespace {
struct CacheKey {
    int dev;
    hipStream_t s;
    bool operator<(const CacheKey& o) const { return dev != o.dev ? dev < o.dev : s < o.s; }
};
struct FreeList {
    std::multimap<size_t, void*> blocks;  // size -> block
    size_t bytes = 0;
};
std::mutex g_cache_mu;
std::map<CacheKey, FreeList> g_free;                         // per stream
std::map<CacheKey, std::pair<void*, size_t>> g_staging;      // per stream pinned buffer
constexpr size_t kCacheCapPerStream = size_t(8) << 30;

size_t round_block(size_t b) {
    size_t r = 256;
    while (r < b) r = r < (1u << 20) ? r * 2 : r + (1u << 20);
    return r;
}

// Frees the cached blocks (and, with staging, the pinned staging buffers) of
// every stream of device `dev`. hipFree waits for the device, so no block is
// freed while queued work may still use it. Caller holds g_cache_mu.
void trim_cache_locked(int dev, bool staging_too) {
    for (auto& kv : g_free) {
        if (kv.first.dev != dev) continue;
        for (auto& b : kv.second.blocks) (void)hipFree(b.second);
        kv.second.blocks.clear();
        kv.second.bytes = 0;
    }
    if (!staging_too) return;
    for (auto it = g_staging.begin(); it != g_staging.end();) {
        if (it->first.dev == dev) {
            if (it->second.first) (void)hipHostFree(it->second.first);
            it = g_staging.erase(it);
        } else {
            ++it;
        }
    }
}
}  // namespace

void* scratch_take(size_t bytes, hipStream_t s, size_t* got) {
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    const size_t want = round_block(bytes);
    {
        std::lock_guard<std::mutex> lock(g_cache_mu);
        FreeList& fl = g_free[CacheKey{dev, s}];
        auto it = fl.blocks.lower_bound(want);
        if (it != fl.blocks.end() && it->first <= 2 * want + (1u << 20)) {
            void* p = it->second;
            *got = it->first;
            fl.bytes -= it->first;
            fl.blocks.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, want);
    if (e == hipErrorOutOfMemory) {  // give the cache back to the runtime and retry once
        (void)hipGetLastError();
        {
            std::lock_guard<std::mutex> lock(g_cache_mu);
            trim_cache_locked(dev, false);
        }
        p = nullptr;
        e = hipMalloc(&p, want);
    }
    D3_HIP(e);
    *got = want;
    return p;
}

void scratch_give(void* p, size_t bytes, hipStream_t s) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return;
    std::lock_guard<std::mutex> lock(g_cache_mu);
    FreeList& fl = g_free[CacheKey{dev, s}];
    if (fl.bytes + bytes > kCacheCapPerStream) {
        (void)hipFree(p);  // waits for the device: the block's last users have finished
        return;
    }
    fl.blocks.emplace(bytes, p);
    fl.bytes += bytes;
}

void release_stream_cache(hipStream_t s) {
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto it = g_free.find(CacheKey{dev, s});
    if (it == g_free.end()) return;
    for (auto& b : it->second.blocks) (void)hipFree(b.second);
    g_free.erase(it);
}

int64_t stream_cache_bytes(hipStream_t s) {
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto it = g_free.find(CacheKey{dev, s});
    return it == g_free.end() ? 0 : (int64_t)it->second.bytes;
}

bool host_is_pinned(const void* p) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // an unregistered pointer is not an error here
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

void release_cache_current_device() {
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    D3_HIP(hipDeviceSynchronize());
    std::lock_guard<std::mutex> lock(g_cache_mu);
    trim_cache_locked(dev, true);
}

static void* staging(size_t bytes, hipStream_t s) {
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_cache_mu);
    auto& st = g_staging[CacheKey{dev, s}];
    if (st.second < bytes) {
        if (st.first) D3_HIP(hipHostFree(st.first));
        st.first = nullptr;
        st.second = 0;
        D3_HIP(hipHostMalloc(&st.first, bytes, hipHostMallocDefault));
        st.second = bytes;
    }
    return st.first;
}

constexpr size_t kStageChunk = size_t(64) << 20;

void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    for (size_t off = 0; off < bytes; off += kStageChunk) {
        const size_t n = std::min(kStageChunk, bytes - off);
        void* st = staging(std::min(kStageChunk, bytes), s);
        std::memcpy(st, static_cast<const char*>(src) + off, n);
        D3_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, st, n, hipMemcpyHostToDevice, s));
        D3_HIP(hipStreamSynchronize(s));
    }
}

void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    for (size_t off = 0; off < bytes; off += kStageChunk) {
        const size_t n = std::min(kStageChunk, bytes - off);
        void* st = staging(std::min(kStageChunk, bytes), s);
        D3_HIP(hipMemcpyAsync(st, static_cast<const char*>(src) + off, n, hipMemcpyDeviceToHost, s));
        D3_HIP(hipStreamSynchronize(s));
        std::memcpy(static_cast<char*>(dst) + off, st, n);
    }
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* __restrict__ p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = v;
}

void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s) {
    if (!n) return;
    const size_t blocks = std::min<size_t>((n + 255) / 256, 4096);
    k_fill_u32<<<(unsigned)blocks, 256, 0, s>>>(p, v, n);
    D3_LAUNCH_CHECK();
}

// A non-blocking stream per thread and device for the synchronous forms.
hipStream_t thread_stream() {
    thread_local std::vector<hipStream_t> streams;
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) D3_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    return streams[dev];
}

static int check_pred_args(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* out) {
    if (n_q < 0 || n_pts < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
    if (n_q > 0 && (!xyz || !op || !idx || !out || n_pts == 0))
        return set_error(DEL3D_ERR_INVALID, "null pointer or empty point set with queries");
    return DEL3D_OK;
}

// Scalar path: evaluated on the host (host_scalar.cpp). The slots' pointer
// identities become ids: a slot gets the id of the first slot holding the same
// pointer (the perturbation compares pointers, src/predicates.cpp:220-227).
static int scalar_query(uint8_t op, const double* const* pts, int np, bool coplanar) {
    P3 p[5];
    int32_t id[5];
    for (int k = 0; k < 5; ++k) {
        const double* v = pts[k < np ? k : np - 1];
        p[k] = P3{v[0], v[1], v[2]};
        id[k] = k;
        for (int j = 0; j < k; ++j)
            if (pts[j < np ? j : np - 1] == v) {
                id[k] = id[j];
                break;
            }
    }
    return coplanar ? host_pred2(op, p, id) : host_pred3(op, p, id);
}

static int scalar_or_abort(uint8_t op, const double* const* pts, int np, bool coplanar = false) {
    int sign = 0;
    const int rc = guarded([&] {
        sign = scalar_query(op, pts, np, coplanar);
        return DEL3D_OK;
    });
    if (rc != DEL3D_OK) {
        std::fprintf(stderr, "del3d_amd: scalar predicate failed: %s\n", last_error_slot().c_str());
        std::abort();
    }
    return sign;
}

}  // namespace d3

using namespace d3;

extern "C" {

const char* del3d_last_error(void) { return last_error_slot().c_str(); }

const char* del3d_version(void) { return "del3d_amd 0.2 (gfx950)"; }

int del3d_release_cache(void) {
    return guarded([&] {
        release_cache_current_device();
        return DEL3D_OK;
    });
}

int del3d_predicates(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                     int64_t n_q, int8_t* sign_out, void* stream) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        launch_predicates(xyz, n_pts, op, idx, n_q, sign_out, (hipStream_t)stream, nullptr);
        return DEL3D_OK;
    });
}

int del3d_predicates_stats(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* sign_out, void* stream, int64_t* stats_host) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        if (!stats_host) return set_error(DEL3D_ERR_INVALID, "stats_host is null");
        hipStream_t s = (hipStream_t)stream;
        Scratch sc(s);
        unsigned long long* st = sc.get<unsigned long long>(kStats);
        fill_u32(reinterpret_cast<uint32_t*>(st), 0u, 2 * kStats, s);
        launch_predicates(xyz, n_pts, op, idx, n_q, sign_out, s, st);
        unsigned long long h[kStats];
        copy_d2h(h, st, sizeof h, s);
        for (int k = 0; k < kStats; ++k) stats_host[k] = (int64_t)h[k];
        return DEL3D_OK;
    });
}

int del3d_predicates_timed(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* sign_out, void* stream, float* ms_out) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        if (!ms_out) return set_error(DEL3D_ERR_INVALID, "ms_out is null");
        if (n_q > (int64_t)(1u << 26))
            return set_error(DEL3D_ERR_INVALID, "timed form takes at most 2^26 queries per call");
        hipStream_t s = (hipStream_t)stream;
        hipEvent_t ev[6] = {};
        for (auto& e : ev) D3_HIP(hipEventCreate(&e));
        struct Destroy {
            hipEvent_t* ev;
            ~Destroy() {
                for (int k = 0; k < 6; ++k)
                    if (ev[k]) (void)hipEventDestroy(ev[k]);
            }
        } destroy{ev};
        if (n_q == 0) {
            ms_out[0] = ms_out[1] = ms_out[2] = 0.0f;
            return DEL3D_OK;
        }
        launch_predicates(xyz, n_pts, op, idx, n_q, sign_out, s, nullptr, ev);
        D3_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < 3; ++k) D3_HIP(hipEventElapsedTime(&ms_out[k], ev[2 * k], ev[2 * k + 1]));
        return DEL3D_OK;
    });
}

int del3d_predicates_ev(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* sign_out, void* stream, void* ev_begin, void* ev_end) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        if (!ev_begin || !ev_end) return set_error(DEL3D_ERR_INVALID, "null event");
        if (n_q > (int64_t)kPredChunk)
            return set_error(DEL3D_ERR_INVALID, "the event form takes at most 2^28 queries per call");
        hipEvent_t ev2[2] = {(hipEvent_t)ev_begin, (hipEvent_t)ev_end};
        if (n_q == 0) {
            D3_HIP(hipEventRecord(ev2[0], (hipStream_t)stream));
            D3_HIP(hipEventRecord(ev2[1], (hipStream_t)stream));
            return DEL3D_OK;
        }
        launch_predicates(xyz, n_pts, op, idx, n_q, sign_out, (hipStream_t)stream, nullptr, nullptr, 0,
                          ev2);
        return DEL3D_OK;
    });
}

int del3d_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                             int64_t n_q, int8_t* sign_out, void* stream) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        if (n_pts > kNoAllocPts)
            return set_error(DEL3D_ERR_INVALID,
                             "del3d_predicates_noalloc takes at most %lld points (its fixed scratch "
                             "holds %zu bytes); use del3d_predicates",
                             (long long)kNoAllocPts, kNoAllocScratchBytes);
        launch_predicates_noalloc(xyz, n_pts, op, idx, n_q, sign_out, (hipStream_t)stream);
        return DEL3D_OK;
    });
}

int del3d_predicates_host(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                          int64_t n_q, int8_t* sign_out) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        if (n_q == 0) return DEL3D_OK;
        hipStream_t s = thread_stream();
        Scratch sc(s);
        double* dx = sc.get<double>((size_t)n_pts * 3);
        uint8_t* dop = sc.get<uint8_t>((size_t)n_q);
        int32_t* did = sc.get<int32_t>((size_t)n_q * 5);
        int8_t* dout = sc.get<int8_t>((size_t)n_q);
        copy_h2d(dx, xyz, (size_t)n_pts * 24, s);
        copy_h2d(dop, op, (size_t)n_q, s);
        copy_h2d(did, idx, (size_t)n_q * 20, s);
        launch_predicates(dx, n_pts, dop, did, n_q, dout, s, nullptr);
        copy_d2h(sign_out, dout, (size_t)n_q, s);
        return DEL3D_OK;
    });
}

int del3d_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                      int64_t n_q, int8_t* out, void* stream) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, out);
        if (rc) return rc;
        launch_predicates2(xyz, n_pts, op, idx, n_q, out, (hipStream_t)stream);
        return DEL3D_OK;
    });
}

int del3d_predicates_mixed(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* out, void* stream) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, out);
        if (rc) return rc;
        hipStream_t s = (hipStream_t)stream;
        // the 3D kernels answer ops 0-2 and write 0 for every other byte; the
        // second pass then answers DEL3D_MIXED_OP2_BASE + k and leaves the
        // 3D queries untouched (stream order puts its writes last)
        launch_predicates(xyz, n_pts, op, idx, n_q, out, s, nullptr);
        launch_predicates2(xyz, n_pts, op, idx, n_q, out, s, DEL3D_MIXED_OP2_BASE);
        return DEL3D_OK;
    });
}

int del3d_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n_cells,
                        int lazy, double* out, void* stream) {
    return guarded([&] {
        if (n_cells < 0 || n_pts < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
        if (n_cells > 0 && (!xyz || !cells || !out || n_pts == 0))
            return set_error(DEL3D_ERR_INVALID, "null pointer or empty point set with cells");
        launch_circumcenters(xyz, n_pts, cells, n_cells, lazy ? 1 : 0, out, (hipStream_t)stream);
        return DEL3D_OK;
    });
}

int del3d_circumcenters_host(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n_cells,
                             int lazy, double* out) {
    return guarded([&] {
        if (n_cells < 0 || n_pts < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
        if (n_cells == 0) return DEL3D_OK;
        if (!xyz || !cells || !out || n_pts == 0)
            return set_error(DEL3D_ERR_INVALID, "null pointer or empty point set with cells");
        hipStream_t s = thread_stream();
        Scratch sc(s);
        double* dx = sc.get<double>((size_t)n_pts * 3);
        int32_t* dc = sc.get<int32_t>((size_t)n_cells * 4);
        double* dout = sc.get<double>((size_t)n_cells * 3);
        copy_h2d(dx, xyz, (size_t)n_pts * 24, s);
        copy_h2d(dc, cells, (size_t)n_cells * 16, s);
        launch_circumcenters(dx, n_pts, dc, n_cells, lazy ? 1 : 0, dout, s);
        copy_d2h(out, dout, (size_t)n_cells * 24, s);
        return DEL3D_OK;
    });
}

int del3d_predicates2_host(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* out) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, out);
        if (rc) return rc;
        if (n_q == 0) return DEL3D_OK;
        hipStream_t s = thread_stream();
        Scratch sc(s);
        double* dx = sc.get<double>((size_t)n_pts * 3);
        uint8_t* dop = sc.get<uint8_t>((size_t)n_q);
        int32_t* did = sc.get<int32_t>((size_t)n_q * 5);
        int8_t* dout = sc.get<int8_t>((size_t)n_q);
        copy_h2d(dx, xyz, (size_t)n_pts * 24, s);
        copy_h2d(dop, op, (size_t)n_q, s);
        copy_h2d(did, idx, (size_t)n_q * 20, s);
        launch_predicates2(dx, n_pts, dop, did, n_q, dout, s);
        copy_d2h(out, dout, (size_t)n_q, s);
        return DEL3D_OK;
    });
}

int del3d_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                       void* stream) {
    return guarded([&] {
        if (n_pts < 0 || n_order < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
        if (n_order > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "more than 2^31-1 indices");
        if (n_order >= 2 && (!xyz || !order || n_pts == 0))
            return set_error(DEL3D_ERR_INVALID, "null pointer or no points");
        spatial_sort_device(xyz, n_pts, order, n_order, (hipStream_t)stream);
        return DEL3D_OK;
    });
}

int del3d_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets, int32_t* order,
                            void* stream) {
    return guarded([&] {
        if (n_sets < 0 || (n_sets > 0 && !off)) return set_error(DEL3D_ERR_INVALID, "bad set table");
        if (n_sets == 0) return DEL3D_OK;
        if (off[0] != 0) return set_error(DEL3D_ERR_INVALID, "off[0] must be 0");
        for (int64_t s = 0; s < n_sets; ++s) {
            if (off[s + 1] < off[s]) return set_error(DEL3D_ERR_INVALID, "offsets decrease at %lld", (long long)s);
            if (off[s + 1] - off[s] > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "set too large");
        }
        if (off[n_sets] > 0 && (!xyz || !order)) return set_error(DEL3D_ERR_INVALID, "null pointer");
        spatial_sort_sets_device(xyz, off, n_sets, order, (hipStream_t)stream);
        return DEL3D_OK;
    });
}

int del3d_spatial_sort_host(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order) {
    return guarded([&] {
        if (n_pts < 0 || n_order < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
        if (n_order > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "more than 2^31-1 indices");
        if (n_order < 2) return DEL3D_OK;
        if (!xyz || !order || n_pts == 0) return set_error(DEL3D_ERR_INVALID, "null pointer or no points");
        hipStream_t s = thread_stream();
        Scratch sc(s);
        double* dx = sc.get<double>((size_t)n_pts * 3);
        int32_t* dord = sc.get<int32_t>((size_t)n_order);
        copy_h2d(dx, xyz, (size_t)n_pts * 24, s);
        copy_h2d(dord, order, (size_t)n_order * 4, s);
        spatial_sort_device(dx, n_pts, dord, n_order, s);
        copy_d2h(order, dord, (size_t)n_order * 4, s);
        return DEL3D_OK;
    });
}

int del3d_orient3d(const double* a, const double* b, const double* c, const double* d) {
    const double* p[4] = {a, b, c, d};
    return scalar_or_abort(DEL3D_OP_ORIENT3D, p, 4);
}

int del3d_in_sphere(const double* a, const double* b, const double* c, const double* d,
                    const double* e) {
    const double* p[5] = {a, b, c, d, e};
    return scalar_or_abort(DEL3D_OP_IN_SPHERE, p, 5);
}

int del3d_in_sphere_unperturbed(const double* a, const double* b, const double* c, const double* d,
                                const double* e) {
    const double* p[5] = {a, b, c, d, e};
    return scalar_or_abort(DEL3D_OP_IN_SPHERE_UNPERTURBED, p, 5);
}

int del3d_orient2d(double px, double py, double qx, double qy, double rx, double ry) {
    const double p[3] = {px, py, 0.0}, q[3] = {qx, qy, 0.0}, r[3] = {rx, ry, 0.0};
    const double* v[3] = {p, q, r};
    return scalar_or_abort(DEL3D_OP2_ORIENT2D_XY, v, 3, true);
}

int del3d_coplanar_orientation(const double* p, const double* q, const double* r) {
    const double* v[3] = {p, q, r};
    return scalar_or_abort(DEL3D_OP2_COPLANAR_ORIENTATION, v, 3, true);
}

int del3d_coplanar_orientation4(const double* p, const double* q, const double* r, const double* s) {
    const double* v[4] = {p, q, r, s};
    return scalar_or_abort(DEL3D_OP2_COPLANAR_ORIENTATION4, v, 4, true);
}

int del3d_collinear(const double* p, const double* q, const double* r) {
    const double* v[3] = {p, q, r};
    return scalar_or_abort(DEL3D_OP2_COLLINEAR, v, 3, true);
}

int del3d_equal(const double* p, const double* q) {
    const double* v[2] = {p, q};
    return scalar_or_abort(DEL3D_OP2_EQUAL, v, 2, true);
}

int del3d_collinear_position(const double* p, const double* q, const double* r) {
    const double* v[3] = {p, q, r};
    return scalar_or_abort(DEL3D_OP2_COLLINEAR_POSITION, v, 3, true);
}

int del3d_coplanar_in_circle(const double* p0, const double* p1, const double* p2, const double* p) {
    const double* v[4] = {p0, p1, p2, p};
    return scalar_or_abort(DEL3D_OP2_IN_CIRCLE, v, 4, true);
}

int del3d_coplanar_in_circle_unperturbed(const double* p0, const double* p1, const double* p2,
                                         const double* p) {
    const double* v[4] = {p0, p1, p2, p};
    return scalar_or_abort(DEL3D_OP2_IN_CIRCLE_UNPERTURBED, v, 4, true);
}

void del3d_circumcenter(const double* p0, const double* p1, const double* p2, const double* p3,
                        int lazy, double* out) {
    const P3 p[4] = {P3{p0[0], p0[1], p0[2]}, P3{p1[0], p1[1], p1[2]}, P3{p2[0], p2[1], p2[2]},
                     P3{p3[0], p3[1], p3[2]}};
    host_circumcenter(p, lazy, out);
}

}  // extern "C"
