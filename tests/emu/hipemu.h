// hipemu.h - runs this repository's HIP kernels on the CPU, for bounds
// checking under AddressSanitizer (test infrastructure only).
//
// A launch runs its blocks one after another; the threads of a block run as
// std::threads, __syncthreads() is a barrier over the block, and each
// wavefront-level operation (__ballot, __shfl, __shfl_up) is a barrier over
// the 64 threads of its wavefront. Kernels must reach wavefront operations
// convergently, which the kernels here do. __shared__ variables become
// function statics: one block runs at a time. Device memory is host memory.
#pragma once

#include <algorithm>
#include <atomic>
#include <barrier>
#include <condition_variable>
#include <mutex>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (&(x))

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};

typedef void* hipStream_t;
typedef void* hipEvent_t;
typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2,
                     hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
enum { hipStreamNonBlocking = 1, hipHostMallocDefault = 0, hipDeviceAttributeMultiprocessorCount = 63 };

namespace emu {
// EMU_FIBERS=1 runs the threads of a block as fibers on the launching OS
// thread (cooperative switches at barriers and wavefront operations): the
// same kernel code, many times faster than one OS thread per lane, but with
// no concurrency for ThreadSanitizer to check.
inline bool fiber_mode() {
    static const bool f = [] {
        const char* e = std::getenv("EMU_FIBERS");
        return e && std::atoi(e) != 0;
    }();
    return f;
}
struct Fiber {
    void* sp = nullptr;
    char* stack = nullptr;
    unsigned tid = 0;
    bool done = false;
};
inline thread_local void* t_sched_sp = nullptr;
inline thread_local Fiber* t_cur = nullptr;
inline void fiber_yield();
// A barrier over `count` participants; in fiber mode a generation counter
// that waiting fibers poll between yields.
struct Bar {
    std::unique_ptr<std::barrier<>> sb;
    unsigned count = 0, arrived = 0;
    uint64_t gen = 0;
    explicit Bar(unsigned n) : count(n) {
        if (!fiber_mode()) sb = std::make_unique<std::barrier<>>(n);
    }
    void arrive_and_wait() {
        if (sb) {
            sb->arrive_and_wait();
            return;
        }
        if (++arrived >= count) {
            arrived = 0;
            ++gen;
            return;
        }
        const uint64_t g = gen;
        while (gen == g) fiber_yield();
    }
    void arrive_and_drop() {
        if (sb) {
            sb->arrive_and_drop();
            return;
        }
        --count;
        if (count > 0 && arrived >= count) {
            arrived = 0;
            ++gen;
        }
    }
};
struct Wave {
    std::unique_ptr<Bar> bar;
    uint64_t vals[64];
    int64_t wide[64];
};
struct Block {
    std::unique_ptr<Bar> bar;
    std::vector<Wave> waves;
    std::atomic<int> and_count{0};  // __syncthreads_and: threads whose predicate was false
    int and_result = 0;
};
inline thread_local dim3 tl_tid, tl_bid;
inline dim3 g_bdim, g_gdim;
inline thread_local Block* tl_block = nullptr;
inline int g_cus = 4;
}  // namespace emu

#define threadIdx (::emu::tl_tid)
#define blockIdx (::emu::tl_bid)
#define blockDim (::emu::g_bdim)
#define gridDim (::emu::g_gdim)

inline void __syncthreads() { emu::tl_block->bar->arrive_and_wait(); }
// barrier, then the AND of every thread's predicate: count the false ones,
// read the count after a second barrier, reset it after a third
inline int __syncthreads_and(int pred) {
    emu::Block* b = emu::tl_block;
    b->bar->arrive_and_wait();
    if (!pred) b->and_count.fetch_add(1);
    b->bar->arrive_and_wait();
    const int r = b->and_count.load() == 0;
    b->bar->arrive_and_wait();
    if (threadIdx.x == 0) b->and_count.store(0);
    b->bar->arrive_and_wait();
    return r;
}
inline int __lane_id() { return (int)(threadIdx.x & 63u); }
inline uint64_t __lanemask_lt() { return (1ull << __lane_id()) - 1ull; }
inline int __popcll(uint64_t v) { return __builtin_popcountll(v); }
inline int __ffsll(long long v) { return __builtin_ffsll(v); }
inline uint32_t __umul24(uint32_t a, uint32_t b) { return (a & 0xFFFFFFu) * (b & 0xFFFFFFu); }

inline emu::Wave& cur_wave() { return emu::tl_block->waves[threadIdx.x >> 6]; }

inline uint64_t __ballot(int pred) {
    emu::Wave& w = cur_wave();
    w.vals[__lane_id()] = pred ? 1 : 0;
    w.bar->arrive_and_wait();
    uint64_t m = 0;
    const unsigned base = threadIdx.x & ~63u;
    for (unsigned l = 0; l < 64 && base + l < emu::g_bdim.x; ++l) m |= (w.vals[l] ? 1ull : 0ull) << l;
    w.bar->arrive_and_wait();
    return m;
}

template <class T>
inline T emu_exchange(T v, int src) {
    emu::Wave& w = cur_wave();
    static_assert(sizeof(T) <= 8, "wide shuffles not emulated");
    int64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    w.wide[__lane_id()] = bits;
    w.bar->arrive_and_wait();
    T r;
    int64_t got = w.wide[src & 63];
    std::memcpy(&r, &got, sizeof(T));
    w.bar->arrive_and_wait();
    return r;
}
template <class T>
inline T __shfl(T v, int src) { return emu_exchange(v, src); }
template <class T>
inline T __shfl_up(T v, unsigned d) {
    const int lane = __lane_id();
    return emu_exchange(v, lane >= (int)d ? lane - (int)d : lane);
}

template <class T>
inline T __shfl_xor(T v, int m) { return emu_exchange(v, __lane_id() ^ m); }

template <class T>
inline T atomicAdd(T* p, T v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicOr(T* p, T v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
template <class T>
inline T atomicMin(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
template <class T>
inline T atomicMax(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {}
    return cur;
}
using std::max;
using std::min;


inline const char* hipGetErrorString(hipError_t) { return "emulated error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
inline hipError_t hipGetSymbolAddress(void** p, const void* sym) {
    *p = const_cast<void*>(sym);
    return hipSuccess;
}
inline hipError_t hipDeviceGetAttribute(int* v, int, int) { *v = emu::g_cus; return hipSuccess; }
namespace emu {
// EMU_DEFER=1 models the runtime's documented asynchrony: kernel launches and
// memsets are queued and run at the next stream synchronization, while
// copies to or from pageable host memory happen at once ("performed
// synchronously" for unpinned memory) and hipFreeAsync releases at once. A
// host sequence that relies on a pageable copy or a free being ordered after
// queued kernels then shows up as a use-after-free under AddressSanitizer.
inline bool defer_mode() {
    static const bool d = [] {
        const char* e = std::getenv("EMU_DEFER");
        return e && std::atoi(e) != 0;
    }();
    return d;
}
inline std::vector<std::function<void()>>& queue() {
    static std::vector<std::function<void()>> q;
    return q;
}
inline std::mutex& queue_mutex() {
    static std::mutex m;
    return m;
}
inline void drain() {
    std::vector<std::function<void()>> q;
    {
        std::lock_guard<std::mutex> lk(queue_mutex());
        q = std::move(queue());
        queue().clear();
    }
    for (auto& f : q) f();
}
inline void enqueue(std::function<void()> f) {
    std::lock_guard<std::mutex> lk(queue_mutex());
    queue().push_back(std::move(f));
}
}  // namespace emu

namespace emu {
// The registry of pinned allocations (several host threads allocate and query).
inline std::vector<std::pair<const char*, size_t>>& pinned() {
    static std::vector<std::pair<const char*, size_t>> v;
    return v;
}
inline std::mutex& pinned_mutex() {
    static std::mutex m;
    return m;
}
inline bool is_pinned(const void* p) {
    std::lock_guard<std::mutex> lk(pinned_mutex());
    for (auto& r : pinned())
        if ((const char*)p >= r.first && (const char*)p < r.first + r.second) return true;
    return false;
}
}  // namespace emu
// Device memory starts with garbage, as on a GPU (large host allocations
// would otherwise read as zeros): every byte is EMU_MALLOC_FILL (default 0xA5).
inline hipError_t hipMallocAsync(void** p, size_t n, hipStream_t) {
    static const int fill = [] {
        const char* e = std::getenv("EMU_MALLOC_FILL");
        return e ? (int)std::strtol(e, nullptr, 0) : 0xA5;
    }();
    *p = std::malloc(n ? n : 1);
    if (*p) std::memset(*p, fill, n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipMalloc(void** p, size_t n) { return hipMallocAsync(p, n, nullptr); }
inline hipError_t hipFreeAsync(void* p, hipStream_t) { std::free(p); return hipSuccess; }
inline hipError_t hipFree(void* p) {
    emu::drain();  // hipFree waits for the device
    std::free(p);
    return hipSuccess;
}
// Copies touching pinned memory are stream-ordered (queued in the deferred
// model); copies of pageable memory happen at once.
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    if (!n) return hipSuccess;
    if (emu::defer_mode() && (emu::is_pinned(d) || emu::is_pinned(s)))
        emu::enqueue([=] { std::memcpy(d, s, n); });
    else
        std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
    if (!n) return hipSuccess;
    if (emu::defer_mode()) emu::enqueue([=] { std::memset(d, v, n); });
    else std::memset(d, v, n);
    return hipSuccess;
}
inline hipError_t hipStreamSynchronize(hipStream_t) {
    emu::drain();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t) {
    emu::drain();
    return hipSuccess;
}
// Streams are distinct handles (the library keys per-stream state on them);
// the emulation orders all work in one queue anyway.
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    static std::atomic<uintptr_t> next{0x1000};
    *s = reinterpret_cast<hipStream_t>(next.fetch_add(0x10));
    return hipSuccess;
}
inline hipError_t hipStreamCreate(hipStream_t* s) { return hipStreamCreateWithFlags(s, 0); }
enum hipMemoryType { hipMemoryTypeUnregistered = 0, hipMemoryTypeHost = 1, hipMemoryTypeDevice = 2 };
struct hipPointerAttribute_t {
    hipMemoryType type;
    int device;
    void* devicePointer;
    void* hostPointer;
    int isManaged;
    unsigned allocationFlags;
};
// Pinned (hipHostMalloc'd) memory is host memory to the runtime; anything
// else is unregistered (an error, as the runtime reports it).
inline hipError_t hipPointerGetAttributes(hipPointerAttribute_t* a, const void* p) {
    if (!emu::is_pinned(p)) return hipErrorInvalidValue;
    a->type = hipMemoryTypeHost;
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() {
    emu::drain();
    return hipSuccess;
}
inline hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
    hipError_t e = hipMallocAsync(p, n, nullptr);
    if (e == hipSuccess) {
        std::lock_guard<std::mutex> lk(emu::pinned_mutex());
        emu::pinned().emplace_back((const char*)*p, n);
    }
    return e;
}
inline hipError_t hipEventCreate(hipEvent_t* e) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipEventDestroy(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t, hipEvent_t) { *ms = 0.0f; return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t);
// Graph capture: the emulation records nothing and replays by running the
// captured sequence again, which is what a graph launch does.
typedef void* hipGraph_t;
typedef void* hipGraphExec_t;
enum { hipStreamCaptureModeThreadLocal = 1 };
inline hipError_t hipStreamBeginCapture(hipStream_t, int) { return hipErrorInvalidValue; }
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t* g) { *g = nullptr; return hipErrorInvalidValue; }
inline hipError_t hipGraphInstantiate(hipGraphExec_t*, hipGraph_t, void*, void*, size_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphLaunch(hipGraphExec_t, hipStream_t) { return hipErrorInvalidValue; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t) { return hipSuccess; }
inline hipError_t hipGraphDestroy(hipGraph_t) { return hipSuccess; }
inline uint32_t __builtin_amdgcn_readfirstlane(uint32_t v) { return v; }
// on the GPU a compiler-only fence for a wavefront's LDS accesses, which the
// hardware executes in order; here the lanes are threads, so a wavefront barrier
inline void __builtin_amdgcn_wave_barrier() {
    emu::Wave& w = cur_wave();
    w.bar->arrive_and_wait();
}
// c + the dot product of the four unsigned bytes of a and b
inline uint32_t __builtin_amdgcn_udot4(uint32_t a, uint32_t b, uint32_t c, bool) {
    for (int i = 0; i < 32; i += 8) c += ((a >> i) & 0xFFu) * ((b >> i) & 0xFFu);
    return c;
}
enum { hipEventDisableTiming = 2 };
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { *e = (void*)1; return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
#define amdgpu_waves_per_eu(...) unused
inline hipError_t hipHostFree(void* p) {
    emu::drain();
    {
        std::lock_guard<std::mutex> lk(emu::pinned_mutex());
        auto& v = emu::pinned();
        for (size_t i = 0; i < v.size(); ++i)
            if (v[i].first == p) { v.erase(v.begin() + i); break; }
    }
    std::free(p);
    return hipSuccess;
}

namespace emu {
template <class K, class... A>
void run_launch(dim3 grid, dim3 block, K kernel, A... args);

template <class K, class... A>
void launch(dim3 grid, dim3 block, K kernel, A... args) {
    if (defer_mode()) {
        enqueue([=] { run_launch(grid, block, kernel, args...); });
        return;
    }
    run_launch(grid, block, kernel, args...);
}

// Persistent workers, one per thread of the largest block seen; each block
// is run by the first block.x workers.
struct Pool {
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv, done_cv;
    uint64_t gen = 0;
    unsigned active = 0, remaining = 0;
    std::function<void(unsigned)> job;
    bool quit = false;
    void ensure(unsigned n) {
        while (workers.size() < n) {
            const unsigned id = (unsigned)workers.size();
            workers.emplace_back([this, id] {
                uint64_t seen = 0;
                for (;;) {
                    std::function<void(unsigned)> j;
                    {
                        std::unique_lock<std::mutex> lk(mu);
                        cv.wait(lk, [&] { return quit || (gen != seen && id < active); });
                        if (quit) return;
                        seen = gen;
                        j = job;
                    }
                    j(id);
                    std::lock_guard<std::mutex> lk(mu);
                    if (--remaining == 0) done_cv.notify_all();
                }
            });
        }
    }
    void run(unsigned n, std::function<void(unsigned)> f) {
        ensure(n);
        std::unique_lock<std::mutex> lk(mu);
        job = std::move(f);
        active = n;
        remaining = n;
        ++gen;
        cv.notify_all();
        done_cv.wait(lk, [&] { return remaining == 0; });
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(mu);
            quit = true;
        }
        cv.notify_all();
        for (auto& w : workers) w.join();
    }
};
inline Pool& pool() {
    static Pool p;
    return p;
}

extern "C" void emu_ctx_switch(void** save_sp, void* load_sp);
asm(R"(
.text
.weak emu_ctx_switch
.type emu_ctx_switch,@function
emu_ctx_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size emu_ctx_switch, .-emu_ctx_switch
)");

inline thread_local std::function<void(unsigned)>* t_job = nullptr;

inline void fiber_yield() { emu_ctx_switch(&t_cur->sp, t_sched_sp); }

[[noreturn]] inline void fiber_entry() {
    Fiber* f = t_cur;
    (*t_job)(f->tid);
    f->done = true;
    emu_ctx_switch(&f->sp, t_sched_sp);
    std::abort();  // a finished fiber is never resumed
}

// Runs job(t) for t in [0, n) as fibers, round robin, until all finish.
inline void run_fibers(unsigned n, std::function<void(unsigned)> job) {
    constexpr size_t kStack = size_t(256) << 10;
    thread_local std::vector<Fiber> fibers;
    if (fibers.size() < n) {
        const size_t old = fibers.size();
        fibers.resize(n);
        for (size_t i = old; i < n; ++i) fibers[i].stack = static_cast<char*>(std::malloc(kStack));
    }
    t_job = &job;
    for (unsigned t = 0; t < n; ++t) {
        Fiber& f = fibers[t];
        f.tid = t;
        f.done = false;
        // initial frame: six callee-saved registers, then fiber_entry as the
        // return address, then a fake caller slot (entry sees rsp % 16 == 8)
        uintptr_t top = (reinterpret_cast<uintptr_t>(f.stack) + kStack) & ~uintptr_t(15);
        void** sp = reinterpret_cast<void**>(top);
        *--sp = nullptr;
        *--sp = reinterpret_cast<void*>(&fiber_entry);
        for (int r = 0; r < 6; ++r) *--sp = nullptr;
        f.sp = sp;
    }
    unsigned alive = n;
    while (alive > 0) {
        for (unsigned t = 0; t < n; ++t) {
            Fiber& f = fibers[t];
            if (f.done) continue;
            t_cur = &f;
            tl_tid = dim3(t);
            emu_ctx_switch(&t_sched_sp, f.sp);
            if (f.done) --alive;
        }
    }
    t_cur = nullptr;
    t_job = nullptr;
}

inline std::mutex& launch_mutex() {
    static std::mutex m;
    return m;
}

// Kernels run one at a time, whatever host thread launches them (__shared__
// variables are statics of the kernel function).
template <class K, class... A>
void run_launch(dim3 grid, dim3 block, K kernel, A... args) {
    std::lock_guard<std::mutex> launch_lock(launch_mutex());
    g_gdim = grid;
    g_bdim = block;
    const unsigned nt = block.x;
    // block order: 0 = launch order, 1 = reversed, 2 = a fixed shuffle; set
    // through EMU_BLOCK_ORDER to expose dependencies between blocks
    static const int order_mode = [] {
        const char* e = std::getenv("EMU_BLOCK_ORDER");
        return e ? std::atoi(e) : 0;
    }();
    std::vector<unsigned> bo(grid.x);
    for (unsigned b = 0; b < grid.x; ++b) bo[b] = b;
    if (order_mode == 1) std::reverse(bo.begin(), bo.end());
    if (order_mode == 2)
        for (unsigned b = grid.x; b > 1; --b) std::swap(bo[b - 1], bo[(b * 2654435761u) % b]);
    for (unsigned bi = 0; bi < grid.x; ++bi) {
        const unsigned b = bo[bi];
        Block blk;
        blk.bar = std::make_unique<Bar>(nt);
        const unsigned nw = (nt + 63) / 64;
        blk.waves.resize(nw);
        for (unsigned w = 0; w < nw; ++w)
            blk.waves[w].bar = std::make_unique<Bar>(std::min(64u, nt - 64 * w));
        if (fiber_mode()) {
            run_fibers(nt, [&, b](unsigned t) {
                tl_bid = dim3(b);
                tl_block = &blk;
                kernel(args...);
                blk.waves[t >> 6].bar->arrive_and_drop();
                blk.bar->arrive_and_drop();
            });
            continue;
        }
        pool().run(nt, [&, b](unsigned t) {
            tl_tid = dim3(t);
            tl_bid = dim3(b);
            tl_block = &blk;
            kernel(args...);
            blk.waves[t >> 6].bar->arrive_and_drop();
            blk.bar->arrive_and_drop();
        });
    }
}
}  // namespace emu
