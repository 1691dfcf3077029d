// capi.hip - the C ABI of include/del3d_amd.h: argument checks, error
// reporting, host-memory and scalar forms over the batched kernels.
#include <mutex>
#include <string>
#include <vector>

#include "common.h"

namespace d3 {

void launch_predicates(const double* xyz, const uint8_t* op, const int32_t* idx, int64_t n_q,
                       int8_t* out, hipStream_t s, unsigned long long* stats_dev);
struct ScalarQuery {
    double xyz[15];
    int32_t id[5];
    uint8_t op;
    int8_t sign;
};
void launch_single(ScalarQuery* qd, hipStream_t s);
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

// A non-blocking stream per thread and device for the synchronous forms.
static hipStream_t thread_stream() {
    thread_local std::vector<hipStream_t> streams;
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    if ((int)streams.size() <= dev) streams.resize(dev + 1, nullptr);
    if (!streams[dev]) D3_HIP(hipStreamCreateWithFlags(&streams[dev], hipStreamNonBlocking));
    return streams[dev];
}

template <class F>
static int guarded(F&& f) {
    try {
        last_error_slot().clear();
        return f();
    } catch (const HipError& e) {
        return e.code;
    } catch (const std::bad_alloc&) {
        return set_error(DEL3D_ERR_NOMEM, "host allocation failed");
    } catch (...) {
        return set_error(DEL3D_ERR_DEVICE, "unexpected exception");
    }
}

static int check_pred_args(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                           int64_t n_q, int8_t* out) {
    if (n_q < 0 || n_pts < 0) return set_error(DEL3D_ERR_INVALID, "negative size");
    if (n_q > 0 && (!xyz || !op || !idx || !out || n_pts == 0))
        return set_error(DEL3D_ERR_INVALID, "null pointer or empty point set with queries");
    return DEL3D_OK;
}

// Scalar path: one pinned staging record and one device record per thread.
struct ScalarSlot {
    ScalarQuery* host = nullptr;
    ScalarQuery* dev = nullptr;
};

static int scalar_query(uint8_t op, const double* const* pts, int np) {
    thread_local std::vector<ScalarSlot> slots;
    int dev = 0;
    D3_HIP(hipGetDevice(&dev));
    if ((int)slots.size() <= dev) slots.resize(dev + 1);
    ScalarSlot& sl = slots[dev];
    if (!sl.host) {
        D3_HIP(hipHostMalloc((void**)&sl.host, sizeof(ScalarQuery), hipHostMallocDefault));
        D3_HIP(hipMalloc((void**)&sl.dev, sizeof(ScalarQuery)));
    }
    ScalarQuery& q = *sl.host;
    q.op = op;
    for (int k = 0; k < 5; ++k) {
        const double* p = pts[k < np ? k : np - 1];
        q.xyz[3 * k] = p[0];
        q.xyz[3 * k + 1] = p[1];
        q.xyz[3 * k + 2] = p[2];
        // pointer identity -> index identity: the first slot holding the same pointer
        int ident = k;
        for (int j = 0; j < k; ++j)
            if (pts[j < np ? j : np - 1] == p) { ident = q.id[j]; break; }
        q.id[k] = ident;
    }
    hipStream_t s = thread_stream();
    D3_HIP(hipMemcpyAsync(sl.dev, sl.host, sizeof(ScalarQuery), hipMemcpyHostToDevice, s));
    launch_single(sl.dev, s);
    D3_HIP(hipMemcpyAsync(sl.host, sl.dev, sizeof(ScalarQuery), hipMemcpyDeviceToHost, s));
    D3_HIP(hipStreamSynchronize(s));
    return q.sign;
}

static int scalar_or_abort(uint8_t op, const double* const* pts, int np) {
    int sign = 0;
    const int rc = guarded([&] {
        sign = scalar_query(op, pts, np);
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

const char* del3d_version(void) { return "del3d_amd 0.1 (gfx950)"; }

int del3d_predicates(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                     int64_t n_q, int8_t* sign_out, void* stream) {
    return guarded([&] {
        const int rc = check_pred_args(xyz, n_pts, op, idx, n_q, sign_out);
        if (rc) return rc;
        launch_predicates(xyz, op, idx, n_q, sign_out, (hipStream_t)stream, nullptr);
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
        DevBuf st(5 * sizeof(unsigned long long), s);
        D3_HIP(hipMemsetAsync(st.as<void>(), 0, 5 * sizeof(unsigned long long), s));
        launch_predicates(xyz, op, idx, n_q, sign_out, s, st.as<unsigned long long>());
        unsigned long long h[5];
        copy_d2h(h, st.as<void>(), sizeof h, s);
        for (int k = 0; k < 5; ++k) stats_host[k] = (int64_t)h[k];
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
        DevBuf dx((size_t)n_pts * 24, s), dop((size_t)n_q, s), did((size_t)n_q * 20, s),
            dout((size_t)n_q, s);
        copy_h2d(dx.as<void>(), xyz, (size_t)n_pts * 24, s);
        copy_h2d(dop.as<void>(), op, (size_t)n_q, s);
        copy_h2d(did.as<void>(), idx, (size_t)n_q * 20, s);
        launch_predicates(dx.as<double>(), dop.as<uint8_t>(), did.as<int32_t>(), n_q,
                          dout.as<int8_t>(), s, nullptr);
        copy_d2h(sign_out, dout.as<void>(), (size_t)n_q, s);
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
        DevBuf dx((size_t)n_pts * 24, s), dord((size_t)n_order * 4, s);
        copy_h2d(dx.as<void>(), xyz, (size_t)n_pts * 24, s);
        copy_h2d(dord.as<void>(), order, (size_t)n_order * 4, s);
        spatial_sort_device(dx.as<double>(), n_pts, dord.as<int32_t>(), n_order, s);
        copy_d2h(order, dord.as<void>(), (size_t)n_order * 4, s);
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

}  // extern "C"
