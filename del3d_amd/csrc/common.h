// common.h - error reporting and HIP call checking for the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/del3d_amd.h"

namespace d3 {

// Message of the calling thread's last failure, returned by del3d_last_error().
std::string& last_error_slot();

inline int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    last_error_slot() = buf;
    return code;
}

struct HipError {
    int code;
};

// Throws HipError after recording the failing call; the C ABI entry points
// catch it and return the code.
#define D3_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            const int c_ = (e_ == hipErrorOutOfMemory) ? DEL3D_ERR_NOMEM : DEL3D_ERR_DEVICE; \
            ::d3::set_error(c_, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                            __FILE__, __LINE__);                                         \
            throw ::d3::HipError{c_};                                                    \
        }                                                                                \
    } while (0)

#define D3_LAUNCH_CHECK() D3_HIP(hipGetLastError())

// Device scratch for one library call, taken from a per-stream cache of
// hipMalloc'd blocks and given back to that cache when the call ends. A block
// given back may still be read by work queued on its stream, so it is only
// handed out again to a later call on the same stream, which the stream
// orders after that work; a block freed to the runtime (cache over its cap,
// out-of-memory retry, del3d_release_cache) goes through hipFree, which waits
// for the device.
void* scratch_take(size_t bytes, hipStream_t s, size_t* got);
void scratch_give(void* p, size_t bytes, hipStream_t s);
// Waits for the current device, then frees every cached scratch block and
// pinned staging buffer of it (del3d_release_cache).
void release_cache_current_device();
// Frees the cached scratch blocks of one stream (hipFree waits for the device).
void release_stream_cache(hipStream_t s);
// Bytes of scratch blocks cached for one stream of the current device.
int64_t stream_cache_bytes(hipStream_t s);
// Whether host memory p is page-locked (hipHostMalloc'd), so that the DMA
// engines can copy it without staging.
bool host_is_pinned(const void* p);

class Scratch {
public:
    explicit Scratch(hipStream_t s) : s_(s) {}
    ~Scratch() {
        for (auto& b : blocks_) scratch_give(b.first, b.second, s_);
    }
    Scratch(const Scratch&) = delete;
    Scratch& operator=(const Scratch&) = delete;
    template <class T>
    T* get(size_t count) {
        size_t got = 0;
        void* p = scratch_take(count * sizeof(T) + 16, s_, &got);
        blocks_.emplace_back(p, got);
        return static_cast<T*>(p);
    }

private:
    hipStream_t s_;
    std::vector<std::pair<void*, size_t>> blocks_;
};

// Host<->device copies go through a pinned staging buffer of the stream, so
// they are ordered in the stream like any kernel; they return when the copy
// has completed.
void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s);
void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s);

// p[0..n) = v, as a kernel on s (no hipMemsetAsync: a kernel is also what a
// captured graph replays reliably).
void fill_u32(uint32_t* p, uint32_t v, size_t n, hipStream_t s);

// Number of compute units of the current device (cached per device).
int device_cu_count();

// A non-blocking stream per thread and device, for the synchronous forms.
hipStream_t thread_stream();

// Runs an entry point's body: clears the thread's error message, turns
// exceptions into status codes.
template <class F>
int guarded(F&& f) {
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

}  // namespace d3
