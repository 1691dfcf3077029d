// common.h - error reporting and HIP call checking for the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
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

// Stream-ordered scratch buffer, released on the same stream.
class DevBuf {
public:
    DevBuf() = default;
    DevBuf(size_t bytes, hipStream_t s) : s_(s) {
        if (bytes) D3_HIP(hipMallocAsync(&p_, bytes, s));
    }
    ~DevBuf() {
        if (p_) (void)hipFreeAsync(p_, s_);
    }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), s_(o.s_) { o.p_ = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            if (p_) (void)hipFreeAsync(p_, s_);
            p_ = o.p_;
            s_ = o.s_;
            o.p_ = nullptr;
        }
        return *this;
    }
    template <class T>
    T* as() const {
        return static_cast<T*>(p_);
    }

private:
    void* p_ = nullptr;
    hipStream_t s_ = nullptr;
};

// Copies between pageable host memory and device memory. The runtime performs
// such copies synchronously, without a promise that queued work on `s` has
// finished (or has not started), so the stream is drained first: a queued
// kernel can neither read what the copy is about to overwrite nor produce,
// after the copy, data the host was meant to read.
inline void copy_h2d(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    D3_HIP(hipStreamSynchronize(s));
    D3_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
    D3_HIP(hipStreamSynchronize(s));
}

inline void copy_d2h(void* dst, const void* src, size_t bytes, hipStream_t s) {
    if (!bytes) return;
    D3_HIP(hipStreamSynchronize(s));
    D3_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, s));
    D3_HIP(hipStreamSynchronize(s));
}

// Scratch for one library call: every allocation lives until the arena is
// destroyed, so no buffer is freed and handed out again while work queued
// earlier in the call may still use it. Destroy it only after the stream has
// been synchronized (or accept that the frees are stream-ordered).
class Scratch {
public:
    explicit Scratch(hipStream_t s) : s_(s) {}
    template <class T>
    T* get(size_t count) {
        bufs_.emplace_back(count * sizeof(T) + 16, s_);
        return bufs_.back().as<T>();
    }

private:
    hipStream_t s_;
    std::vector<DevBuf> bufs_;
};

// Number of compute units of the current device (cached per device).
int device_cu_count();

}  // namespace d3
