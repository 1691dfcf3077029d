// spatial_sort.hip - the reference's spatial sort, bit-exact, on gfx950.
//
// The reference (src/spatial_sort.cpp:141-151) shuffles the indices with a
// rand48-driven Fisher-Yates loop (src/rand48.h:113-153), then runs the
// multiscale pass (:130-137) and, on each multiscale segment, the Hilbert
// median recursion (:95-117) whose median splits are a middle-pivot Lomuto
// quickselect (:40-79). The result depends on the exact draw sequence and on
// the exact arrangement every Lomuto round leaves behind, so the GPU does not
// approximate either; it evaluates them in closed, parallel form.
//
// Shuffle. The accept/reject decision for Fisher-Yates step k depends only on
// k and the draw, never on n, so the steps of a shuffle of n elements are a
// prefix of one universal step sequence j_1, j_2, ... A draw u can only be
// rejected when u >= 2^31 - (K-1), so the few such candidates are compacted
// and resolved by one wavefront (speculative prefix of rejections, iterated to
// its fixed point); every other draw maps to its step in parallel. The final
// position of each element follows from the steps by two relations: F(p), the
// first step after p that selects position p, and G(k), the next step after
// k selecting the same position as k. Element k ends at terminal(F-chain
// from G(k)) if G(k) exists, else at j_k; element 0 at terminal(F(0)) or 0.
// Terminals come from pointer jumping. Truncating F and G at n gives the
// shuffle of any n <= K from the same tables.
//
// Median splits. Per Lomuto round on [L, R] with pivot slot pv: after the
// pivot swap, the elements less than the pivot move, in order, to L, L+1, ...
// An element that is not less, starting at p, is carried along the chain
// q <- Lpos[q] (Lpos[L+r] = position of the r-th smaller element) while
// q < L + nless; the one that ends at L + nless is swapped with the pivot at
// R. So a round is: flags, a per-range prefix, the Lpos table, one scatter.
//   - ranges longer than kWave run these rounds device-wide (tiles of
//     kTileElems records, several ranges batched per launch);
//   - ranges of at most kWave elements finish in one wavefront with keys in
//     LDS; whole Hilbert subtrees of at most kWave elements are sorted by one
//     wavefront from start to finish.
// The recursion plan (split positions, axes, directions) does not depend on
// the data; only the arrangement inside each range does. The host walks the
// plan level by level and batches every range of a phase into one launch.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace d3 {

// ---------------------------------------------------------------------------
// rand48 with jump-ahead (src/rand48.h:26-50)
// ---------------------------------------------------------------------------
constexpr uint64_t kMask48 = (1ull << 48) - 1;
constexpr uint64_t kA48 = 0x5DEECE66Dull;
constexpr uint64_t kC48 = 0xBull;
constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Aff {
    uint64_t a, c;  // x -> a*x + c (mod 2^48)
};

__host__ __device__ inline Aff aff_then(Aff f, Aff g) {  // g after f
    return Aff{(g.a * f.a) & kMask48, (g.a * f.c + g.c) & kMask48};
}

__host__ __device__ inline Aff aff_steps(uint64_t t) {
    Aff r{1, 0}, b{kA48, kC48};
    while (t) {
        if (t & 1) r = aff_then(r, b);
        b = aff_then(b, b);
        t >>= 1;
    }
    return r;
}

constexpr int kDrawsPerThread = 64;

// u[t] = (state after t+1 steps) >> 17, for t in [0, count)
__global__ __launch_bounds__(256) void k_draws(uint32_t* __restrict__ u, uint64_t count,
                                               uint64_t x0) {
    const uint64_t t0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDrawsPerThread;
    if (t0 >= count) return;
    const Aff f = aff_steps(t0 + 1);
    uint64_t x = (f.a * x0 + f.c) & kMask48;
    const uint64_t end = min(count, t0 + kDrawsPerThread);
    for (uint64_t t = t0; t < end; ++t) {
        u[t] = (uint32_t)(x >> 17);
        x = (kA48 * x + kC48) & kMask48;
    }
}

// ---------------------------------------------------------------------------
// Ordered compaction of draw indices with u >= thr (candidates for rejection)
// ---------------------------------------------------------------------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t* s_wave, uint32_t& total) {
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    // inclusive scan within the wavefront
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0;
    total = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        const uint32_t c = s_wave[w];
        before += (w < wave) ? c : 0u;
        total += c;
    }
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(kScanBlock) void k_cand_count(const uint32_t* __restrict__ u,
                                                           uint64_t count, uint32_t thr,
                                                           uint32_t* __restrict__ blk) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanBlock + threadIdx.x;
        c += (t < count && u[t] >= thr) ? 1u : 0u;
    }
    uint32_t total;
    (void)block_exclusive_scan(c, s_wave, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// exclusive scan of a small array in one block (in place), total to *sum
__global__ __launch_bounds__(1024) void k_scan_small(uint32_t* __restrict__ a, uint32_t n,
                                                     uint32_t* __restrict__ sum) {
    __shared__ uint32_t s_wave[16];
    __shared__ uint32_t s_carry;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, total);
        const uint32_t carry = s_carry;
        if (i < n) a[i] = carry + ex;
        __syncthreads();
        if (threadIdx.x == 0) s_carry = carry + total;
        __syncthreads();
    }
    if (threadIdx.x == 0) *sum = s_carry;
}

__global__ __launch_bounds__(kScanBlock) void k_cand_write(const uint32_t* __restrict__ u,
                                                           uint64_t count, uint32_t thr,
                                                           const uint32_t* __restrict__ blk,
                                                           uint32_t* __restrict__ cand) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t off = blk[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanBlock + threadIdx.x;
        const bool f = t < count && u[t] >= thr;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(f ? 1u : 0u, s_wave, total);
        if (f) cand[off + ex] = (uint32_t)t;
        off += total;
    }
}

// One wavefront walks the candidates in draw order. The step of candidate c
// is t_c + 1 - (rejections before it); within a chunk of 64 the number of
// earlier rejections is found by iterating prefix -> decisions -> prefix to
// its fixed point (lane i is final after at most i+1 iterations).
// out[0] = number of rejections; rej[] = rejected draw indices, ascending.
__global__ __launch_bounds__(64) void k_resolve(const uint32_t* __restrict__ u,
                                                const uint32_t* __restrict__ cand,
                                                const uint32_t* __restrict__ ncand_p,
                                                uint32_t last_step, uint32_t* __restrict__ rej,
                                                uint32_t* __restrict__ out) {
    const uint32_t ncand = *ncand_p;
    const int lane = __lane_id();
    const uint64_t lt = __lanemask_lt();
    uint32_t s_in = 0;
    bool finished = false;
    for (uint32_t base = 0; base < ncand && !finished; base += 64) {
        const uint32_t c = base + lane;
        const bool valid = c < ncand;
        const uint32_t t = valid ? cand[c] : 0u;
        const uint32_t uv = valid ? u[t] : 0u;
        uint32_t r = 0;
        uint64_t m = 0;
        bool in_range = false;
        for (int it = 0; it < 65; ++it) {
            const uint32_t k = t + 1u - (s_in + r);
            in_range = valid && k <= last_step;
            bool rj = false;
            if (in_range) {
                const uint32_t b = 0x80000000u / (k + 1u);
                rj = (uv / b) > k;
            }
            m = __ballot(rj);
            const uint32_t rn = (uint32_t)__popcll(m & lt);
            const bool changed = rn != r;
            r = rn;
            if (__ballot(changed) == 0) break;
        }
        const uint32_t k = t + 1u - (s_in + r);
        const bool rj = in_range && ((uv / (0x80000000u / (k + 1u))) > k);
        if (rj) rej[s_in + r] = t;
        s_in += (uint32_t)__popcll(m);
        if (__ballot(valid && !in_range) != 0) finished = true;
    }
    if (lane == 0) out[0] = s_in;
}

__device__ inline uint32_t lower_bound_u32(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// j[k] for steps 1..last_step from the accepted draws
__global__ __launch_bounds__(256) void k_steps(const uint32_t* __restrict__ u, uint64_t count,
                                               const uint32_t* __restrict__ rej,
                                               const uint32_t* __restrict__ nrej_p,
                                               uint32_t last_step, uint32_t* __restrict__ j,
                                               uint32_t* __restrict__ gcount) {
    const uint32_t nrej = *nrej_p;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t lb = lower_bound_u32(rej, nrej, (uint32_t)t);
        if (lb < nrej && rej[lb] == (uint32_t)t) continue;
        const uint64_t k = t + 1 - lb;
        if (k < 1 || k > last_step) continue;
        const uint32_t b = 0x80000000u / (uint32_t)(k + 1);
        const uint32_t jk = u[t] / b;
        j[k] = jk;
        atomicAdd(&gcount[jk], 1u);
    }
}

// block-level tile scan of a large array: per-block sums, then offsets
__global__ __launch_bounds__(kScanBlock) void k_tile_sum(const uint32_t* __restrict__ a, uint64_t n,
                                                         uint32_t* __restrict__ blk) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanBlock + threadIdx.x;
        c += t < n ? a[t] : 0u;
    }
    uint32_t total;
    (void)block_exclusive_scan(c, s_wave, total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanBlock) void k_tile_scan(const uint32_t* __restrict__ a, uint64_t n,
                                                          const uint32_t* __restrict__ blk,
                                                          uint32_t* __restrict__ out) {
    __shared__ uint32_t s_wave[kScanBlock / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t off = blk[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanBlock + threadIdx.x;
        const uint32_t v = t < n ? a[t] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, s_wave, total);
        if (t < n) out[t] = off + ex;
        off += total;
    }
}

// scatter step k into its group (unordered within the group)
__global__ __launch_bounds__(256) void k_group_scatter(const uint32_t* __restrict__ j,
                                                       uint32_t last_step,
                                                       const uint32_t* __restrict__ gstart,
                                                       uint32_t* __restrict__ gfill,
                                                       uint32_t* __restrict__ grp) {
    for (uint64_t k = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= last_step;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = j[k];
        grp[gstart[p] + atomicAdd(&gfill[p], 1u)] = (uint32_t)k;
    }
}

// sort each group ascending; F(p) = first member > p; G(member) = next member
__global__ __launch_bounds__(256) void k_group_links(uint32_t K, const uint32_t* __restrict__ gstart,
                                                     const uint32_t* __restrict__ gcount,
                                                     uint32_t* __restrict__ grp,
                                                     uint32_t* __restrict__ F,
                                                     uint32_t* __restrict__ G) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < K;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = gstart[p], c = gcount[p];
        uint32_t* g = grp + s;
        for (uint32_t i = 1; i < c; ++i) {
            const uint32_t v = g[i];
            uint32_t k = i;
            while (k > 0 && g[k - 1] > v) {
                g[k] = g[k - 1];
                --k;
            }
            g[k] = v;
        }
        uint32_t f = kNone;
        for (uint32_t i = 0; i < c; ++i) {
            if (g[i] > p) {
                f = g[i];
                break;
            }
        }
        F[p] = f;
        for (uint32_t i = 0; i < c; ++i) G[g[i]] = (i + 1 < c) ? g[i + 1] : kNone;
    }
}

// ---------------------------------------------------------------------------
// Per-set part of the shuffle over concatenated sets
// ---------------------------------------------------------------------------
__device__ inline uint32_t set_of(const int64_t* off, uint32_t nsets, uint64_t g) {
    uint32_t lo = 0, hi = nsets;  // last s with off[s] <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// nxt[g] = global position of F(p) if it lies inside the set, else g
__global__ __launch_bounds__(256) void k_chain_init(const int64_t* __restrict__ off, uint32_t nsets,
                                                    uint64_t total, const uint32_t* __restrict__ F,
                                                    uint32_t* __restrict__ nxt) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = set_of(off, nsets, g);
        const uint64_t b = (uint64_t)off[s], n = (uint64_t)off[s + 1] - b;
        const uint32_t p = (uint32_t)(g - b);
        const uint32_t f = F[p];
        nxt[g] = (f != kNone && f < n) ? (uint32_t)(b + f) : (uint32_t)g;
    }
}

__global__ __launch_bounds__(256) void k_jump(const uint32_t* __restrict__ a, uint64_t total,
                                              uint32_t* __restrict__ b) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x)
        b[g] = a[a[g]];
}

// shuffled[final position of k] = order[k]; records are built at once
struct Rec {
    double x, y, z;
    int32_t idx;
    int32_t pad;
};

__global__ __launch_bounds__(256) void k_shuffle_place(const int64_t* __restrict__ off,
                                                       const int64_t* __restrict__ poff,
                                                       uint32_t nsets, uint64_t total,
                                                       const uint32_t* __restrict__ j,
                                                       const uint32_t* __restrict__ F,
                                                       const uint32_t* __restrict__ G,
                                                       const uint32_t* __restrict__ term,
                                                       const int32_t* __restrict__ order,
                                                       const double* __restrict__ xyz,
                                                       Rec* __restrict__ A) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = set_of(off, nsets, g);
        const uint64_t b = (uint64_t)off[s], n = (uint64_t)off[s + 1] - b;
        const uint32_t k = (uint32_t)(g - b);
        uint64_t pos;
        if (n < 2) {
            pos = k;
        } else if (k == 0) {
            const uint32_t f = F[0];
            pos = (f != kNone && f < n) ? (uint64_t)term[b + f] - b : 0;
        } else {
            const uint32_t gk = G[k];
            pos = (gk != kNone && gk < n) ? (uint64_t)term[b + gk] - b : (uint64_t)j[k];
        }
        const int32_t id = order[g];
        const double* p = xyz + 3 * ((uint64_t)poff[s] + (uint64_t)id);
        Rec r;
        r.x = p[0];
        r.y = p[1];
        r.z = p[2];
        r.idx = id;
        r.pad = 0;
        A[b + pos] = r;
    }
}

// ---------------------------------------------------------------------------
// Device-wide Lomuto rounds (ranges longer than kWave)
// ---------------------------------------------------------------------------
constexpr int kWave = 1024;          // ranges up to this size are finished by one wavefront
constexpr int kTileThreads = 256;
constexpr int kTileItems = 16;
constexpr int kTileElems = kTileThreads * kTileItems;

struct SelTask {
    int64_t L, R, pv;  // current inclusive range and this round's pivot slot
    int32_t axis, up;
    int64_t nless;     // output of the round
};

struct SelTile {
    int32_t task;
    int32_t len;
    int64_t start;
    int64_t first_tile;  // index of the task's first tile (for the per-task offset)
};

__device__ inline double rec_key(const Rec& r, int axis) {
    return axis == 0 ? r.x : (axis == 1 ? r.y : r.z);
}

// comparator of src/spatial_sort.cpp:33-35: up ? key(b) < key(a) : key(a) < key(b)
__device__ inline bool key_less(double a, double b, int up) { return up ? (b < a) : (a < b); }

// the arrangement after the pivot swap of src/spatial_sort.cpp:41
__device__ inline const Rec& rec_at(const Rec* A, const SelTask& t, int64_t p) {
    return p == t.pv ? A[t.R] : A[p];
}

__global__ __launch_bounds__(kTileThreads) void k_sel_count(const Rec* __restrict__ A,
                                                            const SelTask* __restrict__ tasks,
                                                            const SelTile* __restrict__ tiles,
                                                            uint32_t* __restrict__ cnt) {
    __shared__ uint32_t s_wave[kTileThreads / 64];
    const SelTile tl = tiles[blockIdx.x];
    const SelTask t = tasks[tl.task];
    const double pk = rec_key(A[t.pv], t.axis);
    uint32_t c = 0;
    for (int i = 0; i < kTileItems; ++i) {
        const int o = i * kTileThreads + threadIdx.x;
        if (o < tl.len) c += key_less(rec_key(rec_at(A, t, tl.start + o), t.axis), pk, t.up) ? 1u : 0u;
    }
    uint32_t total;
    (void)block_exclusive_scan(c, s_wave, total);
    if (threadIdx.x == 0) cnt[blockIdx.x] = total;
}

// per task: exclusive offsets over its tiles, nless
__global__ __launch_bounds__(256) void k_sel_offsets(SelTask* __restrict__ tasks, uint32_t ntasks,
                                                     const int64_t* __restrict__ task_tile0,
                                                     const uint32_t* __restrict__ cnt,
                                                     uint32_t* __restrict__ toff) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < ntasks; i += gridDim.x * blockDim.x) {
        int64_t acc = 0;
        for (int64_t k = task_tile0[i]; k < task_tile0[i + 1]; ++k) {
            toff[k] = (uint32_t)acc;
            acc += cnt[k];
        }
        tasks[i].nless = acc;
    }
}

__global__ __launch_bounds__(kTileThreads) void k_sel_flagged(const Rec* __restrict__ A,
                                                              const SelTask* __restrict__ tasks,
                                                              const SelTile* __restrict__ tiles,
                                                              const uint32_t* __restrict__ toff,
                                                              int64_t* __restrict__ Lpos,
                                                              Rec* __restrict__ B) {
    __shared__ uint32_t s_wave[kTileThreads / 64];
    const SelTile tl = tiles[blockIdx.x];
    const SelTask t = tasks[tl.task];
    const double pk = rec_key(A[t.pv], t.axis);
    int64_t rank = toff[blockIdx.x];
    for (int i = 0; i < kTileItems; ++i) {
        const int o = i * kTileThreads + threadIdx.x;
        const int64_t p = tl.start + o;
        bool f = false;
        Rec r;
        if (o < tl.len) {
            r = rec_at(A, t, p);
            f = key_less(rec_key(r, t.axis), pk, t.up);
        }
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(f ? 1u : 0u, s_wave, total);
        if (f) {
            Lpos[t.L + rank + ex] = p;
            B[t.L + rank + ex] = r;
        }
        rank += total;
    }
}

__global__ __launch_bounds__(kTileThreads) void k_sel_unflagged(const Rec* __restrict__ A,
                                                                const SelTask* __restrict__ tasks,
                                                                const SelTile* __restrict__ tiles,
                                                                const int64_t* __restrict__ Lpos,
                                                                Rec* __restrict__ B) {
    const SelTile tl = tiles[blockIdx.x];
    const SelTask t = tasks[tl.task];
    const double pk = rec_key(A[t.pv], t.axis);
    const int64_t lim = t.L + t.nless;
    for (int i = 0; i < kTileItems; ++i) {
        const int o = i * kTileThreads + threadIdx.x;
        if (o >= tl.len) continue;
        const int64_t p = tl.start + o;
        const Rec r = rec_at(A, t, p);
        if (key_less(rec_key(r, t.axis), pk, t.up)) continue;
        int64_t q = p;
        while (q < lim) q = Lpos[q];
        B[q == lim ? t.R : q] = r;
    }
    if (tl.start == t.L && threadIdx.x == 0) B[lim] = A[t.pv];  // the pivot
}

__global__ __launch_bounds__(kTileThreads) void k_sel_copy(Rec* __restrict__ A,
                                                           const SelTask* __restrict__ tasks,
                                                           const SelTile* __restrict__ tiles,
                                                           const Rec* __restrict__ B) {
    const SelTile tl = tiles[blockIdx.x];
    const SelTask t = tasks[tl.task];
    for (int i = 0; i < kTileItems; ++i) {
        const int o = i * kTileThreads + threadIdx.x;
        if (o < tl.len) A[tl.start + o] = B[tl.start + o];
    }
    if (tl.start == t.L && threadIdx.x == 0) A[t.R] = B[t.R];
}

// ---------------------------------------------------------------------------
// One wavefront per range: quickselect and whole Hilbert subtrees in LDS
// ---------------------------------------------------------------------------
struct WaveTask {
    int64_t L, R, nth;  // inclusive range and target (quickselect finish)
    int32_t axis, up;
};

struct LeafNode {
    int64_t lo, hi;  // [lo, hi)
    int32_t x, ux, uy, uz;
};

// One Lomuto round over slots [l, r] (local) of loc[], comparing key[loc[.]],
// pivot at slot pv. Returns the pivot's final slot. 64 threads (one wavefront).
__device__ inline int wave_partition(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp,
                                     int l, int r, int pv, int up) {
    const int lane = threadIdx.x;
    const uint64_t lt = __lanemask_lt();
    if (lane == 0) {
        const uint16_t v = loc[pv];
        loc[pv] = loc[r];
        loc[r] = v;
    }
    __syncthreads();
    const double pk = key[loc[r]];
    int nless = 0;
    for (int c = l; c < r; c += 64) {
        const int p = c + lane;
        const bool f = p < r && key_less(key[loc[p]], pk, up);
        const uint64_t m = __ballot(f);
        if (f) lp[nless + __popcll(m & lt)] = (uint16_t)p;
        nless += __popcll(m);
    }
    __syncthreads();
    const int lim = l + nless;
    int rank = 0;
    for (int c = l; c < r; c += 64) {
        const int p = c + lane;
        const bool valid = p < r;
        const bool f = valid && key_less(key[loc[p]], pk, up);
        const uint64_t m = __ballot(f);
        if (valid) {
            int dest;
            if (f) {
                dest = l + rank + __popcll(m & lt);
            } else {
                int q = p;
                while (q < lim) q = lp[q - l];
                dest = (q == lim) ? r : q;
            }
            tmp[dest] = loc[p];
        }
        rank += __popcll(m);
    }
    if (lane == 0) tmp[lim] = loc[r];
    __syncthreads();
    for (int p = l + lane; p <= r; p += 64) loc[p] = tmp[p];
    __syncthreads();
    return lim;
}

// quickselect of src/spatial_sort.cpp:60-71 over local slots [l, r]
__device__ inline void wave_select(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp,
                                   int l, int r, int nth, int up) {
    for (;;) {
        const int pv = l + (r - l) / 2;
        const int pos = wave_partition(key, loc, lp, tmp, l, r, pv, up);
        if (pos == nth) return;
        if (nth < pos) r = pos - 1;
        else l = pos + 1;
    }
}

// hilbert_split of src/spatial_sort.cpp:74-79 on local slots [b, e)
__device__ inline int wave_split(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp,
                                 int b, int e, int up) {
    if (b >= e) return b;
    const int mid = b + (e - b) / 2;
    if (e - b >= 2) wave_select(key, loc, lp, tmp, b, e - 1, mid, up);
    return mid;
}

__global__ __launch_bounds__(64) void k_wave_select(Rec* __restrict__ A, Rec* __restrict__ B,
                                                    const WaveTask* __restrict__ tasks,
                                                    uint32_t ntasks) {
    __shared__ double s_key[kWave];
    __shared__ uint16_t s_loc[kWave], s_lp[kWave], s_tmp[kWave];
    for (uint32_t ti = blockIdx.x; ti < ntasks; ti += gridDim.x) {
        const WaveTask t = tasks[ti];
        const int m = (int)(t.R - t.L + 1);
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += 64) {
            s_key[i] = rec_key(A[t.L + i], t.axis);
            s_loc[i] = (uint16_t)i;
        }
        __syncthreads();
        wave_select(s_key, s_loc, s_lp, s_tmp, 0, m - 1, (int)(t.nth - t.L), t.up);
        for (int i = threadIdx.x; i < m; i += 64) B[t.L + i] = A[t.L + s_loc[i]];
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += 64) A[t.L + i] = B[t.L + i];
    }
}

__global__ __launch_bounds__(64) void k_wave_subtree(const Rec* __restrict__ A,
                                                     const LeafNode* __restrict__ nodes,
                                                     uint32_t nnodes, int32_t* __restrict__ order_out) {
    __shared__ double s_key[3][kWave];
    __shared__ uint16_t s_loc[kWave], s_lp[kWave], s_tmp[kWave];
    __shared__ int32_t s_stack[64][6];
    for (uint32_t ni = blockIdx.x; ni < nnodes; ni += gridDim.x) {
        const LeafNode nd = nodes[ni];
        const int m = (int)(nd.hi - nd.lo);
        __syncthreads();
        for (int i = threadIdx.x; i < m; i += 64) {
            const Rec r = A[nd.lo + i];
            s_key[0][i] = r.x;
            s_key[1][i] = r.y;
            s_key[2][i] = r.z;
            s_loc[i] = (uint16_t)i;
        }
        if (threadIdx.x == 0) {
            s_stack[0][0] = 0;
            s_stack[0][1] = m;
            s_stack[0][2] = nd.x;
            s_stack[0][3] = nd.ux;
            s_stack[0][4] = nd.uy;
            s_stack[0][5] = nd.uz;
        }
        __syncthreads();
        int sp = 1;  // uniform across the wavefront
        while (sp > 0) {
            --sp;
            const int b = s_stack[sp][0], e = s_stack[sp][1];
            const int x = s_stack[sp][2], ux = s_stack[sp][3], uy = s_stack[sp][4], uz = s_stack[sp][5];
            __syncthreads();
            if (e - b <= 8) continue;
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            const int m0 = b, m8 = e;
            const int m4 = wave_split(s_key[x], s_loc, s_lp, s_tmp, m0, m8, ux);
            const int m2 = wave_split(s_key[y], s_loc, s_lp, s_tmp, m0, m4, uy);
            const int m1 = wave_split(s_key[z], s_loc, s_lp, s_tmp, m0, m2, uz);
            const int m3 = wave_split(s_key[z], s_loc, s_lp, s_tmp, m2, m4, !uz);
            const int m6 = wave_split(s_key[y], s_loc, s_lp, s_tmp, m4, m8, !uy);
            const int m5 = wave_split(s_key[z], s_loc, s_lp, s_tmp, m4, m6, uz);
            const int m7 = wave_split(s_key[z], s_loc, s_lp, s_tmp, m6, m8, !uz);
            // children of src/spatial_sort.cpp:109-116 (order of processing is
            // irrelevant: the ranges are disjoint)
            const int ch[8][6] = {{m0, m1, z, uz, ux, uy},   {m1, m2, y, uy, uz, ux},
                                  {m2, m3, y, uy, uz, ux},   {m3, m4, x, ux, !uy, !uz},
                                  {m4, m5, x, ux, !uy, !uz}, {m5, m6, y, !uy, uz, !ux},
                                  {m6, m7, y, !uy, uz, !ux}, {m7, m8, z, !uz, !ux, uy}};
            if (threadIdx.x == 0) {
                int w = sp;
                for (int c = 0; c < 8; ++c) {
                    if (ch[c][1] - ch[c][0] <= 8) continue;
                    for (int f = 0; f < 6; ++f) s_stack[w][f] = ch[c][f];
                    ++w;
                }
            }
            int add = 0;
            for (int c = 0; c < 8; ++c) add += (ch[c][1] - ch[c][0] > 8) ? 1 : 0;
            sp += add;
            __syncthreads();
        }
        for (int i = threadIdx.x; i < m; i += 64) order_out[nd.lo + i] = A[nd.lo + s_loc[i]].idx;
    }
}

// ---------------------------------------------------------------------------
// Host driver
// ---------------------------------------------------------------------------
struct Node {
    int64_t lo, hi;
    int x, ux, uy, uz;
};

static inline int64_t mid_of(int64_t b, int64_t e) { return b >= e ? b : b + (e - b) / 2; }

static int grid_for(uint64_t work, int per_block, int cus) {
    const uint64_t blocks = (work + per_block - 1) / per_block;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)cus * 32));
}

// Universal shuffle tables for steps 1..K-1: j, F, G (device arrays of K).
struct ShuffleTables {
    uint32_t *j = nullptr, *F = nullptr, *G = nullptr;
};

static ShuffleTables build_shuffle_tables(uint64_t K, hipStream_t s, Scratch& sc) {
    const int cus = device_cu_count();
    const uint32_t last_step = (uint32_t)(K - 1);
    // expected rejections ~ K^2 / 2^33; generous margin, re-run if exceeded
    const double er = (double)K * (double)K / 8589934592.0;
    uint64_t D = K - 1 + (uint64_t)(er + 12.0 * std::sqrt(er + 1.0)) + 256;
    const uint64_t x0 = ((1ull << 16) | 0x330Eull) & kMask48;  // seed 1
    for (;;) {
        uint32_t* u = sc.get<uint32_t>(D);
        const uint64_t nthreads = (D + kDrawsPerThread - 1) / kDrawsPerThread;
        k_draws<<<(unsigned)((nthreads + 255) / 256), 256, 0, s>>>(u, D, x0);
        D3_LAUNCH_CHECK();
        const uint32_t thr = 0x80000000u - last_step;  // 2^31 - (K-1)
        const uint64_t nblk = (D + kScanTile - 1) / kScanTile;
        uint32_t* blkp = sc.get<uint32_t>(nblk);
        uint32_t* metap = sc.get<uint32_t>(4);  // [0] ncand, [1] nrej, [2] group total
        k_cand_count<<<(unsigned)nblk, kScanBlock, 0, s>>>(u, D, thr, blkp);
        D3_LAUNCH_CHECK();
        k_scan_small<<<1, 1024, 0, s>>>(blkp, (uint32_t)nblk, metap);
        D3_LAUNCH_CHECK();
        uint32_t ncand_h = 0;
        copy_d2h(&ncand_h, metap, 4, s);
        uint32_t* cand = sc.get<uint32_t>((uint64_t)ncand_h + 1);
        uint32_t* rej = sc.get<uint32_t>((uint64_t)ncand_h + 1);
        k_cand_write<<<(unsigned)nblk, kScanBlock, 0, s>>>(u, D, thr, blkp, cand);
        D3_LAUNCH_CHECK();
        k_resolve<<<1, 64, 0, s>>>(u, cand, metap, last_step, rej, metap + 1);
        D3_LAUNCH_CHECK();
        uint32_t nrej_h = 0;
        copy_d2h(&nrej_h, metap + 1, 4, s);
        if (D < (uint64_t)last_step + nrej_h) {  // not enough draws: redo with more
            D = (uint64_t)last_step + nrej_h + 4096 + D / 64;
            continue;
        }
        ShuffleTables T;
        T.j = sc.get<uint32_t>(K);
        T.F = sc.get<uint32_t>(K);
        T.G = sc.get<uint32_t>(K);
        uint32_t* gcount = sc.get<uint32_t>(K);
        uint32_t* gstart = sc.get<uint32_t>(K);
        uint32_t* gfill = sc.get<uint32_t>(K);
        uint32_t* grp = sc.get<uint32_t>(K);
        fill_u32(gcount, 0u, K, s);
        fill_u32(gfill, 0u, K, s);
        fill_u32(T.G, kNone, K, s);
        k_steps<<<grid_for(D, 256, cus), 256, 0, s>>>(u, D, rej, metap + 1, last_step, T.j, gcount);
        D3_LAUNCH_CHECK();
        const uint64_t kb = (K + kScanTile - 1) / kScanTile;
        uint32_t* kblk = sc.get<uint32_t>(kb);
        k_tile_sum<<<(unsigned)kb, kScanBlock, 0, s>>>(gcount, K, kblk);
        D3_LAUNCH_CHECK();
        k_scan_small<<<1, 1024, 0, s>>>(kblk, (uint32_t)kb, metap + 2);
        D3_LAUNCH_CHECK();
        k_tile_scan<<<(unsigned)kb, kScanBlock, 0, s>>>(gcount, K, kblk, gstart);
        D3_LAUNCH_CHECK();
        k_group_scatter<<<grid_for(K, 256, cus), 256, 0, s>>>(T.j, last_step, gstart, gfill, grp);
        D3_LAUNCH_CHECK();
        k_group_links<<<grid_for(K, 256, cus), 256, 0, s>>>((uint32_t)K, gstart, gcount, grp, T.F, T.G);
        D3_LAUNCH_CHECK();
        return T;
    }
}

// Runs the median splits `tasks` (begin, end, axis, up) on A, exactly as
// hilbert_split would, all ranges disjoint.
struct SplitReq {
    int64_t b, e;
    int axis, up;
};

static void run_splits(Rec* A, Rec* B, int64_t* Lpos, const std::vector<SplitReq>& reqs,
                       hipStream_t s, Scratch& sc) {
    const int cus = device_cu_count();
    std::vector<SelTask> big;
    std::vector<int64_t> big_nth;
    std::vector<WaveTask> small;
    for (const SplitReq& r : reqs) {
        if (r.e - r.b < 2) continue;  // empty or single: nth_element returns at once
        const int64_t nth = r.b + (r.e - r.b) / 2;
        if (r.e - r.b <= kWave) {
            small.push_back(WaveTask{r.b, r.e - 1, nth, r.axis, r.up});
        } else {
            big.push_back(SelTask{r.b, r.e - 1, 0, r.axis, r.up, 0});
            big_nth.push_back(nth);
        }
    }
    // device-wide rounds while a range is longer than one wavefront's share
    while (!big.empty()) {
        std::vector<SelTile> tiles;
        std::vector<int64_t> tile0;
        for (size_t i = 0; i < big.size(); ++i) {
            SelTask& t = big[i];
            t.pv = t.L + (t.R - t.L) / 2;
            tile0.push_back((int64_t)tiles.size());
            for (int64_t st = t.L; st < t.R; st += kTileElems) {
                const int64_t len = std::min<int64_t>(kTileElems, t.R - st);
                tiles.push_back(SelTile{(int32_t)i, (int32_t)len, st, tile0.back()});
            }
        }
        tile0.push_back((int64_t)tiles.size());
        SelTask* dt = sc.get<SelTask>(big.size());
        SelTile* dtl = sc.get<SelTile>(tiles.size());
        int64_t* dt0 = sc.get<int64_t>(tile0.size());
        uint32_t* dcnt = sc.get<uint32_t>(tiles.size());
        uint32_t* doff = sc.get<uint32_t>(tiles.size());
        copy_h2d(dt, big.data(), big.size() * sizeof(SelTask), s);
        copy_h2d(dtl, tiles.data(), tiles.size() * sizeof(SelTile), s);
        copy_h2d(dt0, tile0.data(), tile0.size() * 8, s);
        const unsigned nt = (unsigned)tiles.size();
        k_sel_count<<<nt, kTileThreads, 0, s>>>(A, dt, dtl, dcnt);
        D3_LAUNCH_CHECK();
        k_sel_offsets<<<grid_for(big.size(), 256, cus), 256, 0, s>>>(dt, (uint32_t)big.size(), dt0,
                                                                     dcnt, doff);
        D3_LAUNCH_CHECK();
        k_sel_flagged<<<nt, kTileThreads, 0, s>>>(A, dt, dtl, doff, Lpos, B);
        D3_LAUNCH_CHECK();
        k_sel_unflagged<<<nt, kTileThreads, 0, s>>>(A, dt, dtl, Lpos, B);
        D3_LAUNCH_CHECK();
        k_sel_copy<<<nt, kTileThreads, 0, s>>>(A, dt, dtl, B);
        D3_LAUNCH_CHECK();
        std::vector<SelTask> back(big.size());
        copy_d2h(back.data(), dt, big.size() * sizeof(SelTask), s);
        std::vector<SelTask> next;
        std::vector<int64_t> next_nth;
        for (size_t i = 0; i < big.size(); ++i) {
            SelTask t = big[i];
            const int64_t pos = t.L + back[i].nless;
            const int64_t nth = big_nth[i];
            if (pos == nth) continue;
            if (nth < pos) t.R = pos - 1;
            else t.L = pos + 1;
            if (t.R - t.L + 1 <= kWave) {
                small.push_back(WaveTask{t.L, t.R, nth, t.axis, t.up});
            } else {
                next.push_back(t);
                next_nth.push_back(nth);
            }
        }
        big.swap(next);
        big_nth.swap(next_nth);
    }
    if (!small.empty()) {
        WaveTask* dw = sc.get<WaveTask>(small.size());
        copy_h2d(dw, small.data(), small.size() * sizeof(WaveTask), s);
        const unsigned g = (unsigned)std::min<size_t>(small.size(), (size_t)cus * 8);
        k_wave_select<<<g, 64, 0, s>>>(A, B, dw, (uint32_t)small.size());
        D3_LAUNCH_CHECK();
    }
}

// Sorts every set: order[off[i] .. off[i+1]) holds indices into the set's
// points, which start at xyz + 3*poff[i] (poff == nullptr: poff = off).
static void sort_sets(const double* xyz, const int64_t* poff, const int64_t* off, int64_t n_sets,
                      int32_t* order, hipStream_t s) {
    const uint64_t total = (uint64_t)off[n_sets];
    if (total == 0) return;
    const int cus = device_cu_count();
    uint64_t K = 2;
    for (int64_t i = 0; i < n_sets; ++i) K = std::max<uint64_t>(K, (uint64_t)(off[i + 1] - off[i]));

    // every buffer of this call lives until the end of the call (see Scratch)
    Scratch sc(s);
    const ShuffleTables T = build_shuffle_tables(K, s, sc);

    int64_t* doff = sc.get<int64_t>(n_sets + 1);
    int64_t* dpoff = sc.get<int64_t>(n_sets + 1);
    copy_h2d(doff, off, (n_sets + 1) * 8, s);
    copy_h2d(dpoff, poff ? poff : off, (n_sets + 1) * 8, s);
    uint32_t* ta = sc.get<uint32_t>(total);
    uint32_t* tb = sc.get<uint32_t>(total);
    k_chain_init<<<grid_for(total, 256, cus), 256, 0, s>>>(doff, (uint32_t)n_sets, total, T.F, ta);
    D3_LAUNCH_CHECK();
    int rounds = 1;
    while ((1ull << rounds) < K) ++rounds;
    uint32_t* cur = ta;
    uint32_t* nxt = tb;
    for (int r = 0; r < rounds; ++r) {
        k_jump<<<grid_for(total, 256, cus), 256, 0, s>>>(cur, total, nxt);
        D3_LAUNCH_CHECK();
        std::swap(cur, nxt);
    }
    Rec* A = sc.get<Rec>(total);
    Rec* B = sc.get<Rec>(total);
    int64_t* lpos = sc.get<int64_t>(total);
    k_shuffle_place<<<grid_for(total, 256, cus), 256, 0, s>>>(doff, dpoff, (uint32_t)n_sets, total,
                                                              T.j, T.F, T.G, cur, order, xyz, A);
    D3_LAUNCH_CHECK();

    // multiscale segments (src/spatial_sort.cpp:130-137) of every set
    std::vector<Node> level, leaves;
    for (int64_t si = 0; si < n_sets; ++si) {
        int64_t b = off[si], e = off[si + 1];
        if (e - b < 2) {
            if (e - b == 1) leaves.push_back(Node{b, e, 0, 0, 0, 0});
            continue;
        }
        for (;;) {
            int64_t mid = b;
            if (e - b >= 64) mid = b + (int64_t)((double)(e - b) * 0.125);
            level.push_back(Node{mid, e, 0, 0, 0, 0});
            if (mid == b) break;
            e = mid;
        }
    }
    // level-synchronous Hilbert recursion for ranges longer than a wavefront's share
    while (!level.empty()) {
        std::vector<Node> big;
        for (const Node& nd : level) {
            if (nd.hi - nd.lo > kWave) big.push_back(nd);
            else leaves.push_back(nd);
        }
        if (big.empty()) break;
        std::vector<SplitReq> px, py, pz;
        for (const Node& nd : big) px.push_back(SplitReq{nd.lo, nd.hi, nd.x, nd.ux});
        run_splits(A, B, lpos, px, s, sc);
        for (const Node& nd : big) {
            const int y = (nd.x + 1) % 3;
            const int64_t m0 = nd.lo, m8 = nd.hi, m4 = mid_of(m0, m8);
            py.push_back(SplitReq{m0, m4, y, nd.uy});
            py.push_back(SplitReq{m4, m8, y, !nd.uy});
        }
        run_splits(A, B, lpos, py, s, sc);
        std::vector<Node> next;
        for (const Node& nd : big) {
            const int x = nd.x, y = (x + 1) % 3, z = (x + 2) % 3;
            const int ux = nd.ux, uy = nd.uy, uz = nd.uz;
            const int64_t m0 = nd.lo, m8 = nd.hi, m4 = mid_of(m0, m8);
            const int64_t m2 = mid_of(m0, m4), m6 = mid_of(m4, m8);
            const int64_t m1 = mid_of(m0, m2), m3 = mid_of(m2, m4);
            const int64_t m5 = mid_of(m4, m6), m7 = mid_of(m6, m8);
            pz.push_back(SplitReq{m0, m2, z, uz});
            pz.push_back(SplitReq{m2, m4, z, !uz});
            pz.push_back(SplitReq{m4, m6, z, uz});
            pz.push_back(SplitReq{m6, m8, z, !uz});
            next.push_back(Node{m0, m1, z, uz, ux, uy});
            next.push_back(Node{m1, m2, y, uy, uz, ux});
            next.push_back(Node{m2, m3, y, uy, uz, ux});
            next.push_back(Node{m3, m4, x, ux, !uy, !uz});
            next.push_back(Node{m4, m5, x, ux, !uy, !uz});
            next.push_back(Node{m5, m6, y, !uy, uz, !ux});
            next.push_back(Node{m6, m7, y, !uy, uz, !ux});
            next.push_back(Node{m7, m8, z, !uz, !ux, uy});
        }
        run_splits(A, B, lpos, pz, s, sc);
        level.swap(next);
    }
    std::vector<LeafNode> lv;
    lv.reserve(leaves.size());
    for (const Node& nd : leaves)
        if (nd.hi > nd.lo) lv.push_back(LeafNode{nd.lo, nd.hi, nd.x, nd.ux, nd.uy, nd.uz});
    if (!lv.empty()) {
        LeafNode* dl = sc.get<LeafNode>(lv.size());
        copy_h2d(dl, lv.data(), lv.size() * sizeof(LeafNode), s);
        const unsigned g = (unsigned)std::min<size_t>(lv.size(), (size_t)cus * 8);
        k_wave_subtree<<<g, 64, 0, s>>>(A, dl, (uint32_t)lv.size(), order);
        D3_LAUNCH_CHECK();
    }
    // the scratch arena is released after this: nothing queued may still use it
    D3_HIP(hipStreamSynchronize(s));
}

void spatial_sort_sets_device(const double* xyz, const int64_t* off, int64_t n_sets,
                              int32_t* order, hipStream_t s) {
    sort_sets(xyz, nullptr, off, n_sets, order, s);
}

// The reference's (points, order) pair: n_order indices into n_pts points.
void spatial_sort_device(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                         hipStream_t s) {
    (void)n_pts;
    if (n_order < 2) return;
    const int64_t off[2] = {0, n_order};
    const int64_t poff[2] = {0, n_pts};
    sort_sets(xyz, poff, off, 1, order, s);
}

}  // namespace d3
