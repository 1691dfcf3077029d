// spatial_sort.hip - the reference's spatial sort, bit-exact, on gfx950.
//
// The reference (src/spatial_sort.cpp:141-151) shuffles the indices with a
// rand48-driven Fisher-Yates loop (src/rand48.h:113-153), then runs the
// multiscale pass (:130-137) and, on each multiscale segment, the Hilbert
// median recursion (:95-117) whose median splits are a middle-pivot Lomuto
// quickselect (:40-79). The permutation depends on the exact draw sequence
// and on the exact arrangement every Lomuto round leaves behind, so neither
// is approximated: both are evaluated in closed, parallel form.
//
// Shuffle. Fisher-Yates step k (k = 1 .. n-1) swaps a[k] with a[j_k],
// j_k in [0, k]. Whether a draw is rejected depends on k and the draw only,
// so the steps of a shuffle of n are a prefix of one universal sequence
// j_1, j_2, ... Only a draw u >= 2^31 - (K-1) can be rejected; those few
// candidates are compacted in draw order and resolved by one wavefront, and
// every other draw maps to its step in parallel. Element k of the input ends
// at j_k if no later step selects j_k, otherwise it follows the chain
// G(k) -> F(.) -> F(.) ..., where G(k) is the next step selecting the same
// position as step k and F(p) the first step after p selecting position p;
// element 0 follows F from 0. Truncating F and G at n gives the shuffle of
// any n <= K from the same tables.
//
// Median splits. A Lomuto round on [L, R] with pivot slot pv swaps a[pv]
// with a[R], then moves the elements less than the pivot, in order, to L,
// L+1, ...; the elements that are not less form a window that starts at the
// next free front slot, and each later smaller element swaps with the
// window's front. The pivot ends at lim = L + (number of smaller elements)
// and what would have landed on lim goes to R. Two closed forms of one round
// are used. Device-wide, destination-driven: a slot in [lim, R) keeps an
// element that is not less and otherwise gathers from the backward chain
// s <- L + rank(s), which needs only the flag bits (every write goes to
// consecutive slots). In LDS, source-driven: an element that is not less,
// at p, follows the forward chain q <- Lpos[q - L] (Lpos[r] = slot of the
// r-th smaller element) while q < lim. The recursion plan (every range, axis
// and direction) does not depend on the data, so the host builds it once and
// uploads it in one copy; only the arrangement inside each range is data.
//   - a select on more than kFin elements runs device-wide rounds: tiles of
//     kTile elements, records in structure-of-arrays form ping-ponged between
//     two buffers, round control kept on the device (the host only reads a
//     counter of still-active selects every few rounds);
//   - a select on at most kFin elements (or a big one once its range has
//     shrunk to that) is finished by one 256-thread workgroup in LDS;
//   - a Hilbert subtree of at most kSub elements is sorted from start to
//     finish by one wavefront in LDS.
//
// Every global-memory index in these kernels is checked against the size of
// the buffer it addresses, and every loop has a bound; a failed check skips
// the access and sets a bit in a device error word, which the call returns
// as DEL3D_ERR_INTERNAL. Device memory outside the call's buffers is never
// touched, whatever a table holds.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

namespace d3 {
namespace {

// ---------------------------------------------------------------------------
// error word
// ---------------------------------------------------------------------------
enum : uint32_t {
    kErrShuffle = 1u,   // a shuffle table entry out of range
    kErrIndex = 2u,     // an order[] entry outside the set's points
    kErrRound = 4u,     // a device-round index out of range
    kErrLds = 8u,       // an LDS-pass range or index out of range
    kErrLoop = 16u,     // a loop hit its bound
};

__device__ inline void raise_err(uint32_t* err, uint32_t bit) { atomicOr(err, bit); }

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

__host__ __device__ inline Aff aff_compose(Aff f, Aff g) {  // g after f
    return Aff{(g.a * f.a) & kMask48, (g.a * f.c + g.c) & kMask48};
}

__host__ __device__ inline Aff aff_pow(uint64_t t) {  // t steps of the generator
    Aff r{1, 0}, b{kA48, kC48};
    while (t) {
        if (t & 1) r = aff_compose(r, b);
        b = aff_compose(b, b);
        t >>= 1;
    }
    return r;
}

constexpr int kDrawRun = 64;

// u[t] = (state after t+1 steps) >> 17, t in [0, count)
__global__ __launch_bounds__(256) void k_draws(uint32_t* __restrict__ u, uint64_t count, uint64_t x0) {
    const uint64_t t0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kDrawRun;
    if (t0 >= count) return;
    const Aff f = aff_pow(t0 + 1);
    uint64_t x = (f.a * x0 + f.c) & kMask48;
    const uint64_t end = t0 + kDrawRun < count ? t0 + kDrawRun : count;
    for (uint64_t t = t0; t < end; ++t) {
        u[t] = (uint32_t)(x >> 17);
        x = (kA48 * x + kC48) & kMask48;
    }
}

// ---------------------------------------------------------------------------
// scans
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive prefix of v over the workgroup; *total = the workgroup's sum.
// Every thread of the workgroup must call it.
__device__ inline uint32_t wg_exclusive(uint32_t v, uint32_t* s_wave, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) s_wave[wave] = x;
    __syncthreads();
    uint32_t before = 0, all = 0;
    const int nw = (int)(blockDim.x >> 6);
    for (int w = 0; w < nw; ++w) {
        const uint32_t c = s_wave[w];
        before += w < wave ? c : 0u;
        all += c;
    }
    __syncthreads();
    *total = all;
    return before + x - v;
}

// blk[b] = number of u[t] >= thr in tile b
__global__ __launch_bounds__(kScanThreads) void k_cand_count(const uint32_t* __restrict__ u, uint64_t count,
                                                             uint32_t thr, uint32_t* __restrict__ blk) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanThreads + threadIdx.x;
        c += (t < count && u[t] >= thr) ? 1u : 0u;
    }
    uint32_t total;
    (void)wg_exclusive(c, s_wave, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// in-place exclusive scan of a[0..n) by one 1024-thread workgroup; *sum = total
__global__ __launch_bounds__(1024) void k_scan_one(uint32_t* __restrict__ a, uint32_t n,
                                                   uint32_t* __restrict__ sum) {
    __shared__ uint32_t s_wave[16];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n; base += blockDim.x) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < n ? a[i] : 0u;
        uint32_t total;
        const uint32_t ex = wg_exclusive(v, s_wave, &total);
        if (i < n) a[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) *sum = carry;
}

// draw indices with u >= thr, in draw order, into cand[0 .. cap)
__global__ __launch_bounds__(kScanThreads) void k_cand_write(const uint32_t* __restrict__ u, uint64_t count,
                                                             uint32_t thr, const uint32_t* __restrict__ blk,
                                                             uint32_t* __restrict__ cand,
                                                             uint32_t* __restrict__ cand_u, uint32_t cap,
                                                             uint32_t* __restrict__ err) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t off = blk[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanThreads + threadIdx.x;
        const uint32_t ut = t < count ? u[t] : 0u;
        const bool f = t < count && ut >= thr;
        uint32_t total;
        const uint32_t ex = wg_exclusive(f ? 1u : 0u, s_wave, &total);
        if (f) {
            if (off + ex < cap) {
                cand[off + ex] = (uint32_t)t;
                cand_u[off + ex] = ut;
            } else {
                raise_err(err, kErrShuffle);
            }
        }
        off += total;
    }
}

// One wavefront walks the candidates in draw order. Candidate c belongs to
// step t_c + 1 - (rejections before it); inside a chunk of 64 the number of
// earlier rejections is found by iterating prefix -> decisions -> prefix to
// its fixed point (lane i is final after at most i+1 iterations).
// out[0] = number of rejections; rej[] = rejected draw indices, ascending.
// The candidates and their draws (cand, cand_u) are staged through LDS in
// blocks of kResBlock, so the walk waits for memory once per block rather
// than once per chunk of 64.
constexpr uint32_t kResBlock = 2048;
__global__ __launch_bounds__(64) void k_resolve(const uint32_t* __restrict__ cand_u, uint64_t count,
                                                const uint32_t* __restrict__ cand,
                                                const uint32_t* __restrict__ ncand_p, uint32_t cap,
                                                uint32_t last_step, uint32_t* __restrict__ rej,
                                                uint32_t* __restrict__ out, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_t[kResBlock], s_u[kResBlock];
    uint32_t ncand = *ncand_p;
    if (ncand > cap) {
        if (threadIdx.x == 0) raise_err(err, kErrShuffle);
        ncand = cap;
    }
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint32_t s_in = 0;
    bool finished = false;
    for (uint32_t blk = 0; blk < ncand && !finished; blk += kResBlock) {
        const uint32_t nb = min(kResBlock, ncand - blk);
        __syncthreads();  // the previous block's reads are done
        for (uint32_t i = (uint32_t)lane; i < nb; i += 64) {
            s_t[i] = cand[blk + i];
            s_u[i] = cand_u[blk + i];
        }
        __syncthreads();
        for (uint32_t base = 0; base < nb && !finished; base += 64) {
            const uint32_t c = base + lane;
            const bool valid = c < nb;
            const uint32_t t = valid ? s_t[c] : 0u;
            const bool t_ok = t < count;
            const uint32_t uv = (valid && t_ok) ? s_u[c] : 0u;
            uint32_t r = 0;
            uint64_t m = 0;
            bool in_range = false;
            for (int it = 0; it < 66; ++it) {
                const uint32_t k = t + 1u - (s_in + r);
                in_range = valid && t_ok && k <= last_step;
                bool rj = false;
                if (in_range) rj = (uv / (0x80000000u / (k + 1u))) > k;
                m = __ballot(rj);
                const uint32_t rn = (uint32_t)__popcll(m & lt);
                const bool changed = rn != r;
                r = rn;
                if (__ballot(changed) == 0) break;
            }
            if ((m >> lane) & 1ull) {
                if (s_in + r < cap) rej[s_in + r] = t;
                else raise_err(err, kErrShuffle);
            }
            s_in += (uint32_t)__popcll(m);
            if (__ballot(valid && !in_range) != 0) finished = true;
        }
    }
    if (lane == 0) out[0] = s_in;
}

__device__ inline uint32_t count_below(const uint32_t* a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;  // number of a[i] < v, a ascending
    while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// j[k] for steps 1..last_step from the accepted draws; gcount[p] = #{k : j_k = p}
__global__ __launch_bounds__(256) void k_steps(const uint32_t* __restrict__ u, uint64_t count,
                                               const uint32_t* __restrict__ rej,
                                               const uint32_t* __restrict__ nrej_p, uint32_t cap,
                                               uint32_t last_step, uint32_t K, uint32_t* __restrict__ j,
                                               uint32_t* __restrict__ gcount, uint32_t* __restrict__ err) {
    uint32_t nrej = *nrej_p;
    if (nrej > cap) nrej = cap;  // k_resolve raised the error
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < count;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t before = count_below(rej, nrej, (uint32_t)t);
        if (before < nrej && rej[before] == (uint32_t)t) continue;  // rejected draw
        const uint64_t k = t + 1 - before;
        if (k < 1 || k > last_step) continue;
        const uint32_t jk = u[t] / (0x80000000u / (uint32_t)(k + 1));
        if (jk > k || k >= K) {
            raise_err(err, kErrShuffle);
            continue;
        }
        j[k] = jk;
        atomicAdd(&gcount[jk], 1u);
    }
}

// per-tile sums of a[0..n) (tiles of kScanTile)
__global__ __launch_bounds__(kScanThreads) void k_tile_sum(const uint32_t* __restrict__ a, uint64_t n,
                                                           uint32_t* __restrict__ blk) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t c = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanThreads + threadIdx.x;
        c += t < n ? a[t] : 0u;
    }
    uint32_t total;
    (void)wg_exclusive(c, s_wave, &total);
    if (threadIdx.x == 0) blk[blockIdx.x] = total;
}

// out = exclusive prefix of a, given the exclusive prefix of the tile sums
__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const uint32_t* __restrict__ a, uint64_t n,
                                                            const uint32_t* __restrict__ blk,
                                                            uint32_t* __restrict__ out) {
    __shared__ uint32_t s_wave[kScanThreads / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
    uint32_t off = blk[blockIdx.x];
    for (int i = 0; i < kScanItems; ++i) {
        const uint64_t t = base + (uint64_t)i * kScanThreads + threadIdx.x;
        const uint32_t v = t < n ? a[t] : 0u;
        uint32_t total;
        const uint32_t ex = wg_exclusive(v, s_wave, &total);
        if (t < n) out[t] = off + ex;
        off += total;
    }
}

// step k into its group's slots (unordered within the group)
__global__ __launch_bounds__(256) void k_group_scatter(const uint32_t* __restrict__ j, uint32_t last_step,
                                                       uint32_t K, const uint32_t* __restrict__ gstart,
                                                       uint32_t* __restrict__ gfill,
                                                       uint32_t* __restrict__ grp,
                                                       uint32_t* __restrict__ err) {
    for (uint64_t k = 1 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= last_step;
         k += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t p = j[k];
        if (p >= K) {
            raise_err(err, kErrShuffle);
            continue;
        }
        const uint64_t slot = (uint64_t)gstart[p] + atomicAdd(&gfill[p], 1u);
        if (slot < K) grp[slot] = (uint32_t)k;
        else raise_err(err, kErrShuffle);
    }
}

// sorts each group ascending; F(p) = first member > p; G(member) = next member
__global__ __launch_bounds__(256) void k_group_links(uint32_t K, const uint32_t* __restrict__ gstart,
                                                     const uint32_t* __restrict__ gcount,
                                                     uint32_t* __restrict__ grp, uint32_t* __restrict__ F,
                                                     uint32_t* __restrict__ G, uint32_t* __restrict__ err) {
    for (uint64_t p = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; p < K;
         p += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = gstart[p], c = gcount[p];
        if ((uint64_t)s + c > K) {
            raise_err(err, kErrShuffle);
            F[p] = kNone;
            continue;
        }
        uint32_t* g = grp + s;
        for (uint32_t i = 1; i < c; ++i) {  // insertion sort; groups hold ~ln(K/p) steps
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
        for (uint32_t i = 0; i < c; ++i) {
            if (g[i] < K) G[g[i]] = (i + 1 < c) ? g[i + 1] : kNone;
            else raise_err(err, kErrShuffle);
        }
    }
}

// ---------------------------------------------------------------------------
// Records: structure of arrays, two buffers (0 = "A", 1 = "B")
// ---------------------------------------------------------------------------
struct Bufs {
    double* x[2];
    double* y[2];
    double* z[2];
    int32_t* id[2];
    int64_t n;  // elements in each array
};

// Buffer selection by a runtime parity is written as selects, not array
// indexing, so the pointers stay in registers.
__device__ inline double* bx(const Bufs& b, int par) { return par ? b.x[1] : b.x[0]; }
__device__ inline double* by(const Bufs& b, int par) { return par ? b.y[1] : b.y[0]; }
__device__ inline double* bz(const Bufs& b, int par) { return par ? b.z[1] : b.z[0]; }
__device__ inline int32_t* bid(const Bufs& b, int par) { return par ? b.id[1] : b.id[0]; }

__device__ inline const double* axis_of(const Bufs& b, int par, int axis) {
    return axis == 0 ? bx(b, par) : (axis == 1 ? by(b, par) : bz(b, par));
}

// comparator of src/spatial_sort.cpp:33-35: up ? key(b) < key(a) : key(a) < key(b)
__device__ inline bool key_less(double a, double b, int up) { return up ? (b < a) : (a < b); }

__device__ inline uint32_t set_of(const int64_t* off, uint32_t nsets, uint64_t g) {
    uint32_t lo = 0, hi = nsets;  // last s with off[s] <= g
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint64_t)off[mid] <= g) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Final position of element k of a set of n >= 2 elements under the shuffle.
__device__ inline uint64_t shuffle_dest(uint32_t k, uint64_t n, const uint32_t* j, const uint32_t* F,
                                        const uint32_t* G, uint32_t K, uint32_t* err) {
    uint32_t start;
    if (k == 0) {
        start = F[0];
        if (start == kNone || start >= n) return 0;
    } else {
        const uint32_t gk = G[k];
        if (gk == kNone || gk >= n) return j[k];
        start = gk;
    }
    uint32_t q = start;
    for (uint64_t hop = 0; hop < n; ++hop) {  // F strictly increases along the chain
        if (q >= K) {
            raise_err(err, kErrShuffle);
            return 0;
        }
        const uint32_t f = F[q];
        if (f == kNone || f >= n) return q;
        q = f;
    }
    raise_err(err, kErrLoop);
    return 0;
}

// Shuffles each set's order[] and gathers the points into buffer 0.
__global__ __launch_bounds__(256) void k_place(const int64_t* __restrict__ off, const int64_t* __restrict__ poff,
                                               uint32_t nsets, uint64_t total, int64_t n_pts,
                                               const uint32_t* __restrict__ j, const uint32_t* __restrict__ F,
                                               const uint32_t* __restrict__ G, uint32_t K,
                                               const int32_t* __restrict__ order,
                                               const double* __restrict__ xyz, Bufs bufs,
                                               uint32_t* __restrict__ err) {
    for (uint64_t g = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
         g += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t s = set_of(off, nsets, g);
        const uint64_t b = (uint64_t)off[s], n = (uint64_t)off[s + 1] - b;
        const uint64_t k = g - b;
        if (k >= n || n > K) {
            raise_err(err, kErrShuffle);
            continue;
        }
        const uint64_t pos = n < 2 ? k : shuffle_dest((uint32_t)k, n, j, F, G, K, err);
        if (pos >= n) {
            raise_err(err, kErrShuffle);
            continue;
        }
        const int64_t id = order[g];
        const int64_t pb = poff[s], pn = poff[s + 1] - pb;
        const uint64_t dst = b + pos;
        if (id < 0 || id >= pn || pb + id >= n_pts || dst >= (uint64_t)bufs.n) {
            raise_err(err, kErrIndex);
            continue;
        }
        const double* p = xyz + 3 * (pb + id);
        bufs.x[0][dst] = p[0];
        bufs.y[0][dst] = p[1];
        bufs.z[0][dst] = p[2];
        bufs.id[0][dst] = (int32_t)id;
    }
}

// ---------------------------------------------------------------------------
// Device-wide Lomuto rounds for selects on more than kFin elements
// ---------------------------------------------------------------------------
constexpr int kFin = 4096;          // selects up to this size finish in LDS
constexpr int kTileThreads = 256;
constexpr int kTileIters = 16;
constexpr int kTile = kTileThreads * kTileIters;  // 4096 elements, 64 mask words
constexpr int kTileWords = kTile / 64;
constexpr int kRing = 4;            // active-select counters, indexed by round

struct RTask {           // a big select of one phase (host-built)
    int64_t b, e;        // original range [b, e)
    int64_t nth;
    int64_t tile0, ntiles;
    int32_t axis, up;
};

struct RState {          // per task, double-buffered by round parity
    int64_t L, R;        // current inclusive range
    int64_t lim;         // L + number of smaller elements, from k_rd_scan
    int64_t m0lo, m0hi;  // margins [lo, hi) finalized by the last round
    int64_t m1lo, m1hi;
    int32_t par;         // buffer holding [L, R]
    int32_t active;      // device rounds continue
    int32_t fin;         // [L, R] is left for the LDS finish
    int32_t mpar;        // the margins are in buffer 1 and are copied to buffer 0
};

struct RTile {
    int64_t start, len;
    int32_t task, pad;
};

__device__ inline bool rd_range_ok(const RState& S, int64_t n) {
    return S.L >= 0 && S.L < S.R && S.R < n;
}

__device__ inline void copy_rec(const Bufs& b, int from, int to, int64_t src, int64_t dst) {
    bx(b, to)[dst] = bx(b, from)[src];
    by(b, to)[dst] = by(b, from)[src];
    bz(b, to)[dst] = bz(b, from)[src];
    bid(b, to)[dst] = bid(b, from)[src];
}

__global__ __launch_bounds__(256) void k_rd_init(const RTask* __restrict__ tasks, uint32_t ntasks,
                                                 RState* __restrict__ st) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntasks; t += gridDim.x * blockDim.x) {
        const RTask T = tasks[t];
        RState S;
        S.L = T.b;
        S.R = T.e - 1;
        S.lim = T.b;
        S.m0lo = S.m0hi = S.m1lo = S.m1hi = 0;
        S.par = 0;
        S.active = 1;
        S.fin = 0;
        S.mpar = 0;
        st[t] = S;
    }
}

// The tiles a round works on (one thread per tile of the phase): those that
// intersect the current range [L, R] of an active select, appended in any
// order to list[] (count in *cnt, zeroed by the previous round's k_rd_count).
// The ranges only shrink, so a round's list is contained in the previous
// one's, which covers the margins that round finalized as well.
__global__ __launch_bounds__(256) void k_rd_list(const RTile* __restrict__ tiles, uint32_t ntiles,
                                                 uint32_t ntasks, const RState* __restrict__ st,
                                                 uint32_t* __restrict__ list, uint32_t* __restrict__ cnt) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        const RTile tl = tiles[t];
        if (tl.task < 0 || (uint32_t)tl.task >= ntasks) continue;  // k_rd_count raises it
        const RState S = st[tl.task];
        if (!S.active || tl.start > S.R || tl.start + tl.len <= S.L) continue;
        const uint32_t k = atomicAdd(cnt, 1u);
        if (k < ntiles) list[k] = t;
    }
}

// Round part 1 (per listed tile, grid-stride): copies the margins the
// previous round left in buffer 1, then flags the tile's elements that are
// less than the pivot (one bit each), counts them, and records each
// 64-element word's exclusive prefix inside the tile. Wavefront w takes
// words 16w .. 16w+15. Block 0 also clears the next round's counters.
__global__ __launch_bounds__(kTileThreads) void k_rd_count(Bufs bufs, const RTask* __restrict__ tasks,
                                                           uint32_t ntasks, const RTile* __restrict__ tiles,
                                                           uint32_t ntiles, const RState* __restrict__ st,
                                                           const uint32_t* __restrict__ list,
                                                           const uint32_t* __restrict__ list_cnt,
                                                           uint32_t* __restrict__ next_list_cnt,
                                                           uint64_t* __restrict__ mask,
                                                           uint16_t* __restrict__ wloc,
                                                           uint32_t* __restrict__ tcount,
                                                           uint32_t* __restrict__ ring, uint32_t round,
                                                           uint32_t* __restrict__ err) {
    constexpr int kWordsPerWave = kTileWords / (kTileThreads / 64);
    __shared__ uint32_t s_cnt[kTileThreads / 64];
    __shared__ uint16_t s_wl[kTileWords];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ring[(round + 1) % kRing] = 0;
        if (next_list_cnt) *next_list_cnt = 0;
    }
    const uint32_t n = min(*list_cnt, ntiles);
    for (uint32_t li = blockIdx.x; li < n; li += gridDim.x) {
        __syncthreads();  // s_cnt / s_wl of the previous tile are read
        const uint32_t tile = list[li];
        if (tile >= ntiles) {
            if (threadIdx.x == 0) raise_err(err, kErrRound);
            continue;
        }
        const RTile tl = tiles[tile];
        if (tl.task < 0 || (uint32_t)tl.task >= ntasks || tl.len < 0 || tl.len > kTile || tl.start < 0 ||
            tl.start + tl.len > bufs.n) {
            if (threadIdx.x == 0) raise_err(err, kErrRound);
            continue;
        }
        const RTask T = tasks[tl.task];
        const RState S = st[tl.task];
        const int64_t t0 = tl.start, t1 = tl.start + tl.len;
        if (S.mpar) {
            for (int m = 0; m < 2; ++m) {
                const int64_t mlo = m ? S.m1lo : S.m0lo, mhi = m ? S.m1hi : S.m0hi;
                const int64_t lo = mlo > t0 ? mlo : t0, hi = mhi < t1 ? mhi : t1;
                if (lo < hi && (lo < 0 || hi > bufs.n)) {
                    if (threadIdx.x == 0) raise_err(err, kErrRound);
                    continue;
                }
                for (int64_t p = lo + threadIdx.x; p < hi; p += kTileThreads) copy_rec(bufs, 1, 0, p, p);
            }
        }
        if (!S.active) {
            if (threadIdx.x == 0) tcount[tile] = 0;
            continue;
        }
        if (!rd_range_ok(S, bufs.n) || (S.par & ~1)) {
            if (threadIdx.x == 0) {
                raise_err(err, kErrRound);
                tcount[tile] = 0;
            }
            continue;
        }
        const double* key = axis_of(bufs, S.par, T.axis);
        const int64_t pv = S.L + (S.R - S.L) / 2;
        const double pk = key[pv];
        const double kr = key[S.R];
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        // every key of the thread's words first, so their loads are in flight
        // together (the mask stores below could otherwise order them)
        double kv[kWordsPerWave];
#pragma unroll
        for (int it = 0; it < kWordsPerWave; ++it) {
            const int64_t p = t0 + (int64_t)(wave * kWordsPerWave + it) * 64 + lane;
            kv[it] = (p < t1 && p >= S.L && p < S.R) ? (p == pv ? kr : key[p]) : 0.0;
        }
        uint32_t c = 0;
#pragma unroll
        for (int it = 0; it < kWordsPerWave; ++it) {
            const int w = wave * kWordsPerWave + it;
            const int64_t p = t0 + (int64_t)w * 64 + lane;
            const bool f = p < t1 && p >= S.L && p < S.R && key_less(kv[it], pk, T.up);
            const uint64_t m = __ballot(f);
            if (lane == 0) {
                mask[(uint64_t)tile * kTileWords + w] = m;
                s_wl[w] = (uint16_t)c;
            }
            c += (uint32_t)__popcll(m);
        }
        if (lane == 0) s_cnt[wave] = c;
        __syncthreads();
        if (threadIdx.x < kTileWords) {
            const int w = threadIdx.x, ww = w / kWordsPerWave;
            uint32_t before = 0;
            for (int i = 0; i < ww; ++i) before += s_cnt[i];
            wloc[(uint64_t)tile * kTileWords + w] = (uint16_t)(before + s_wl[w]);
        }
        if (threadIdx.x == 0) {
            uint32_t sum = 0;
            for (int i = 0; i < kTileThreads / 64; ++i) sum += s_cnt[i];
            tcount[tile] = sum;
        }
    }
}

// Round part 2 (per task): exclusive prefix of the counts of the task's tiles
// that intersect [L, R], and lim. An inactive task's state is carried to the
// next round's buffer here (with its margins, copied by this round's
// k_rd_count, cleared), since no block of k_rd_move handles it.
__global__ __launch_bounds__(256) void k_rd_scan(const RTask* __restrict__ tasks, uint32_t ntasks,
                                                 uint32_t ntiles, RState* __restrict__ st,
                                                 RState* __restrict__ st_next,
                                                 const uint32_t* __restrict__ tcount,
                                                 uint32_t* __restrict__ toff, uint32_t* __restrict__ err) {
    __shared__ uint32_t s_wave[4];
    const uint32_t t = blockIdx.x;
    if (t >= ntasks) return;
    const RTask T = tasks[t];
    const RState S = st[t];
    if (!S.active) {
        if (threadIdx.x == 0) {
            RState N = S;
            N.mpar = 0;
            N.m0lo = N.m0hi = N.m1lo = N.m1hi = 0;
            st_next[t] = N;
        }
        return;
    }
    if (T.tile0 < 0 || T.ntiles < 0 || T.tile0 + T.ntiles > (int64_t)ntiles || S.L < T.b || S.R >= T.e ||
        S.L > S.R) {
        if (threadIdx.x == 0) raise_err(err, kErrRound);
        return;
    }
    const int64_t tb = (S.L - T.b) / kTile, te = (S.R - T.b) / kTile + 1;  // the tiles of [L, R]
    uint32_t carry = 0;
    for (int64_t base = tb; base < te; base += blockDim.x) {
        const int64_t i = base + threadIdx.x;
        const uint32_t v = i < te ? tcount[T.tile0 + i] : 0u;
        uint32_t total;
        const uint32_t ex = wg_exclusive(v, s_wave, &total);
        if (i < te) toff[T.tile0 + i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) st[t].lim = S.L + (int64_t)carry;
}

// Round part 3 (one slot per thread, kSlices blocks per tile): the slots of
// the elements that are not less than the pivot, in order: the j-th of them
// (from 0) at npos[L + j], as an offset from the task's first slot. The move
// kernel's chains jump with it (below).
constexpr int kSlices = kTile / kTileThreads;

__global__ __launch_bounds__(kTileThreads) void k_rd_npos(const RTask* __restrict__ tasks, uint32_t ntasks,
                                                          const RTile* __restrict__ tiles, uint32_t ntiles,
                                                          const RState* __restrict__ st,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ list_cnt,
                                                          const uint64_t* __restrict__ mask,
                                                          const uint16_t* __restrict__ wloc,
                                                          const uint32_t* __restrict__ toff,
                                                          uint32_t* __restrict__ npos, int64_t n) {
    const uint32_t units = min(*list_cnt, ntiles) * kSlices;
    for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
        const uint32_t tile = list[u / kSlices], slice = u % kSlices;
        if (tile >= ntiles) continue;  // k_rd_count raised it
        const RTile tl = tiles[tile];
        if (tl.task < 0 || (uint32_t)tl.task >= ntasks) continue;  // k_rd_count raised it
        const RTask T = tasks[tl.task];
        const RState S = st[tl.task];
        if (!S.active || !rd_range_ok(S, n) || S.L < T.b || S.R >= T.e) continue;  // k_rd_move raises it
        const int lane = threadIdx.x & 63;
        const int w = (int)slice * (kTileThreads / 64) + (threadIdx.x >> 6);
        const int64_t p = tl.start + (int64_t)w * 64 + lane;
        if (p >= tl.start + tl.len || p < S.L || p >= S.R) continue;
        const uint64_t gw = (uint64_t)tile * kTileWords + w;
        const uint64_t m = mask[gw];
        if ((m >> lane) & 1ull) continue;
        const int64_t rank = (int64_t)toff[tile] + wloc[gw] + __popcll(m & ((1ull << lane) - 1ull));
        const int64_t j = (p - S.L) - rank;  // not-less elements in [L, p)
        if (j >= 0 && S.L + j < S.R) npos[S.L + j] = (uint32_t)(p - T.b);
    }
}

// Round part 4 (one slot per thread, kSlices blocks per tile): the whole
// Lomuto round as a permutation written into the other buffer.
//   - an element less than the pivot, at p, goes to L + rank(p), where
//     rank(p) = number of such elements in [L, p) (order kept);
//   - a slot d in [lim, R) whose element is not less keeps it (the window of
//     not-less elements only ever starts below lim, so such an element is
//     never at its front);
//   - a slot d in [lim, R) whose element is less receives the element at the
//     window's front when the scan reached d, the slot L + rank(d); if the
//     element that slot started with is less as well, it had itself taken
//     the front of its own time, so s <- L + rank(s) is followed (strictly
//     down) until it reaches a slot whose element is not less;
//   - what lands on lim goes to R, and the pivot to lim
//     (src/spatial_sort.cpp:41-50).
// rank() needs only the flag words and their prefixes, which are small and
// stay in cache; every write of this kernel is to consecutive slots.
// A chain is long when nearly every element is less than the pivot. Between
// two consecutive not-less slots P = npos[L + j - 1] < s, every slot is less
// and has j not-less slots below it, so each hop there is s <- s - j: after
// a few single hops a chain jumps a whole such interval at once, to
// s - j * ceil((s - P) / j), which is P itself (the chain ends on the element
// there) or a slot below P with fewer not-less slots below it.

__global__ __launch_bounds__(kTileThreads) void k_rd_move(Bufs bufs, const RTask* __restrict__ tasks,
                                                          uint32_t ntasks, const RTile* __restrict__ tiles,
                                                          uint32_t ntiles, const RState* __restrict__ st,
                                                          RState* __restrict__ st_next,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ list_cnt,
                                                          const uint64_t* __restrict__ mask,
                                                          const uint16_t* __restrict__ wloc,
                                                          const uint32_t* __restrict__ toff,
                                                          const uint32_t* __restrict__ npos,
                                                          uint32_t* __restrict__ ring, uint32_t round,
                                                          uint32_t* __restrict__ err) {
  const uint32_t units = min(*list_cnt, ntiles) * kSlices;
  for (uint32_t u = blockIdx.x; u < units; u += gridDim.x) {
    const uint32_t tile = list[u / kSlices], slice = u % kSlices;
    if (tile >= ntiles) continue;  // k_rd_count raised it
    const RTile tl = tiles[tile];
    if (tl.task < 0 || (uint32_t)tl.task >= ntasks) continue;  // k_rd_count raised it
    const RTask T = tasks[tl.task];
    const RState S = st[tl.task];
    if (!S.active) continue;  // listed tiles belong to active selects; k_rd_scan carries the others
    const bool ok = rd_range_ok(S, bufs.n) && !(S.par & ~1) && S.lim >= S.L && S.lim <= S.R &&
                    T.tile0 >= 0 && T.tile0 + T.ntiles <= (int64_t)ntiles && S.L >= T.b && S.R < T.e;
    // the block of the tile holding L, slice 0, finishes the round for the
    // select: the pivot's move and the next state
    const bool first = ok && slice == 0 && (int64_t)tile == T.tile0 + (S.L - T.b) / kTile;
    if (!ok) {
        if (threadIdx.x == 0) raise_err(err, kErrRound);
        continue;
    }
    const int64_t pv = S.L + (S.R - S.L) / 2;
    {
        const int src = S.par, dst = S.par ^ 1;
        const int lane = threadIdx.x & 63;
        const int w = (int)slice * (kTileThreads / 64) + (threadIdx.x >> 6);
        const int64_t p = tl.start + (int64_t)w * 64 + lane;
        const uint64_t gw = (uint64_t)tile * kTileWords + w;
        const uint64_t m = mask[gw];
        if (p < tl.start + tl.len && p >= S.L && p < S.R) {
            const bool f = (m >> lane) & 1ull;
            if (f) {  // less: to the front, in order
                const int64_t d = S.L + (int64_t)toff[tile] + wloc[gw] + __popcll(m & ((1ull << lane) - 1ull));
                if (d >= S.L && d < S.lim) copy_rec(bufs, src, dst, p == pv ? S.R : p, d);
                else raise_err(err, kErrRound);
            }
            if (p >= S.lim) {
                int64_t s = p;
                bool less = f;
                // the three words rank(s) needs, carried so a hop costs one load latency
                uint64_t ms = m;
                uint32_t ws = wloc[gw], ts_off = toff[tile];
                int64_t os = p - T.b;
                const int64_t span = S.R - S.L;
                int64_t hops = 0;
                bool bad = false;
                while (less) {
                    if (++hops > span) {
                        bad = true;
                        break;
                    }
                    const int64_t rank = (int64_t)ts_off + ws + __popcll(ms & ((1ull << (os & 63)) - 1ull));
                    if (hops <= 3) {
                        s = S.L + rank;
                    } else {
                        const int64_t j = (s - S.L) - rank;  // not-less slots below s
                        if (j < 1 || j > span) {
                            bad = true;
                            break;
                        }
                        const int64_t P = T.b + (int64_t)npos[S.L + j - 1];
                        if (P < S.L || P >= s) {
                            bad = true;
                            break;
                        }
                        s -= j * ((s - P + j - 1) / j);
                        if (s == P) break;  // the not-less element at P
                    }
                    if (s < S.L || s >= S.lim) {
                        bad = true;
                        break;
                    }
                    os = s - T.b;
                    const int64_t tsi = T.tile0 + os / kTile;
                    const uint64_t gws = (uint64_t)tsi * kTileWords + (uint64_t)((os % kTile) >> 6);
                    ms = mask[gws];
                    ws = wloc[gws];
                    ts_off = toff[tsi];
                    less = (ms >> (os & 63)) & 1ull;
                }
                if (bad) raise_err(err, hops > span ? kErrLoop : kErrRound);
                else copy_rec(bufs, src, dst, s == pv ? S.R : s, p == S.lim ? S.R : p);
            }
        }
        if (first && threadIdx.x == 0) copy_rec(bufs, src, dst, pv, S.lim);  // the pivot
    }
    if (!(first && threadIdx.x == 0)) continue;
    RState N = S;
    N.mpar = 0;
    N.m0lo = N.m0hi = N.m1lo = N.m1hi = 0;
    if (ok) {
        const int dst = S.par ^ 1;
        const int64_t pos = S.lim;
        N.par = dst;
        N.mpar = dst;  // margins are final in the buffer just written
        N.active = 0;
        N.fin = 0;
        if (pos == T.nth) {
            N.m0lo = S.L;
            N.m0hi = S.R + 1;
            N.L = N.R = pos;
        } else {
            int64_t L = S.L, R = S.R;
            if (T.nth < pos) R = pos - 1;
            else L = pos + 1;
            N.m0lo = S.L;
            N.m0hi = L;
            N.m1lo = R + 1;
            N.m1hi = S.R + 1;
            N.L = L;
            N.R = R;
            const int64_t size = R - L + 1;
            if (size > kFin) {
                N.active = 1;
                atomicAdd(&ring[(round + 1) % kRing], 1u);
            } else if (size >= 2) {
                N.fin = 1;
            } else {  // one slot: the element at nth is in place; all of it is final
                N.m0lo = S.L;
                N.m0hi = S.R + 1;
                N.m1lo = N.m1hi = 0;
            }
        }
    }
    st_next[tl.task] = N;
  }
}

// ---------------------------------------------------------------------------
// LDS passes: Lomuto rounds over local slots, keys read through loc[]
// ---------------------------------------------------------------------------
// One round over local slots [l, r] with pivot slot pv, by a workgroup of NW
// wavefronts; returns the pivot's final slot. loc[] holds the arrangement
// (indices into key[]); lp[] and tmp[] are scratch of the same size.
template <int NW>
__device__ int lds_partition(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp,
                             uint32_t* s_cnt, int l, int r, int pv, int up, uint32_t* err) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    if (tid == 0) {
        const uint16_t v = loc[pv];
        loc[pv] = loc[r];
        loc[r] = v;
    }
    __syncthreads();
    const double pk = key[loc[r]];
    const int span = r - l;
    const int cw = (span + NW - 1) / NW;
    const int wb = l + w * cw;
    const int we = wb + cw < r ? wb + cw : r;
    int cnt = 0;
    for (int c = wb; c < we; c += 64) {
        const int p = c + lane;
        const bool f = p < we && key_less(key[loc[p]], pk, up);
        cnt += __popcll(__ballot(f));
    }
    if (lane == 0) s_cnt[w] = (uint32_t)cnt;
    __syncthreads();
    int before = 0, nless = 0;
    for (int i = 0; i < NW; ++i) {
        const int v = (int)s_cnt[i];
        before += i < w ? v : 0;
        nless += v;
    }
    const int lim = l + nless;
    int rank = before;
    for (int c = wb; c < we; c += 64) {
        const int p = c + lane;
        const bool f = p < we && key_less(key[loc[p]], pk, up);
        const uint64_t m = __ballot(f);
        if (f) {
            const int rk = rank + __popcll(m & lt);
            lp[rk] = (uint16_t)p;
            tmp[l + rk] = loc[p];
        }
        rank += __popcll(m);
    }
    __syncthreads();
    for (int c = wb; c < we; c += 64) {
        const int p = c + lane;
        if (p >= we || key_less(key[loc[p]], pk, up)) continue;
        int q = p, hops = 0;
        while (q < lim && hops <= span) {
            q = lp[q - l];
            ++hops;
        }
        if (hops > span || (q != lim && (q < lim || q > r))) {
            raise_err(err, kErrLds);
            continue;
        }
        tmp[q == lim ? r : q] = loc[p];
    }
    if (tid == 0) tmp[lim] = loc[r];
    __syncthreads();
    for (int p = l + tid; p <= r; p += NW * 64) loc[p] = tmp[p];
    __syncthreads();
    return lim;
}

// Position of the t-th (from 0) set bit of m; m has more than t set bits.
__device__ inline int select_bit(uint64_t m, int t) {
    int lo = 0, hi = 63;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (__popcll(m & ((2ull << mid) - 1ull)) >= t + 1) hi = mid;
        else lo = mid + 1;
    }
    return lo;
}

// The rounds of a select on local slots [l, r] of at most 64 elements, in one
// wavefront's registers: lane i holds slot l + i (its loc[] entry and key).
// A round is the same Lomuto round as lds_partition, evaluated by pulls:
// slot d < lim takes the (d - l)-th smaller element; slot lim the pivot;
// a slot in (lim, r) keeps an element that is not less and otherwise takes
// the end of the backward chain s <- l + rank(s) (found by pointer jumping:
// the chain is shorter than 64 steps, so six doublings reach its end); slot
// r takes the end of the chain from lim. Called by the 64 lanes of one
// wavefront; loc[l .. r] is rewritten at the end.
__device__ void wave_select(const double* key, uint16_t* loc, int l, int r, int nth, int up, uint32_t* err) {
    const int lane = __lane_id();
    const int m = r - l + 1;
    if (m < 1 || m > 64 || nth < l || nth > r) {
        if (lane == 0) raise_err(err, kErrLds);
        return;
    }
    const bool in = lane < m;
    uint32_t id = in ? loc[l + lane] : 0u;
    double k = in ? key[id] : 0.0;
    int a = 0, b = m - 1;
    const int t = nth - l;
    bool done = false;
    for (int round = 0; round <= m && !done; ++round) {
        const int pv = a + (b - a) / 2;
        // the pivot to slot b
        const uint32_t id_pv = __shfl(id, pv), id_b = __shfl(id, b);
        const double k_pv = __shfl(k, pv), k_b = __shfl(k, b);
        if (lane == b) {
            id = id_pv;
            k = k_pv;
        } else if (lane == pv) {
            id = id_b;
            k = k_b;
        }
        const bool f = lane >= a && lane < b && key_less(k, k_pv, up);
        const uint64_t mask = __ballot(f);
        const int rank = __popcll(mask & ((1ull << lane) - 1ull));
        const int lim = a + __popcll(mask);
        int h = f ? a + rank : lane;  // backward chain; fixed at slots that are not less
        for (int j = 0; j < 6; ++j) h = __shfl(h, h);
        const int h_lim = __shfl(h, lim);
        int src = lane;
        if (lane >= a && lane <= b) {
            if (lane < lim) src = select_bit(mask, lane - a);
            else if (lane == lim) src = b;
            else if (lane == b) src = h_lim;
            else src = h;
        }
        id = __shfl(id, src);
        k = __shfl(k, src);
        if (lim == t) {
            done = true;
        } else {
            if (t < lim) b = lim - 1;
            else a = lim + 1;
            if (a > b) break;
        }
    }
    if (!done && lane == 0) raise_err(err, kErrLoop);
    if (in) loc[l + lane] = (uint16_t)id;
}

// quickselect of src/spatial_sort.cpp:60-71 over local slots [l, r]; once the
// range holds at most 64 elements, the first wavefront finishes it in
// registers
template <int NW>
__device__ void lds_select(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp, uint32_t* s_cnt,
                           int l, int r, int nth, int up, uint32_t* err) {
    const int bound = r - l + 1;  // every round removes at least the pivot's slot
    for (int round = 0; round <= bound; ++round) {
        if (r - l < 64) {
            if (threadIdx.x < 64) wave_select(key, loc, l, r, nth, up, err);
            __syncthreads();
            return;
        }
        const int pv = l + (r - l) / 2;
        const int pos = lds_partition<NW>(key, loc, lp, tmp, s_cnt, l, r, pv, up, err);
        if (pos == nth) return;
        if (nth < pos) r = pos - 1;
        else l = pos + 1;
        if (l > r) break;
    }
    if (threadIdx.x == 0) raise_err(err, kErrLoop);
}

// hilbert_split of src/spatial_sort.cpp:74-79 on local slots [b, e)
template <int NW>
__device__ int lds_split(const double* key, uint16_t* loc, uint16_t* lp, uint16_t* tmp, uint32_t* s_cnt,
                         int b, int e, int up, uint32_t* err) {
    if (b >= e) return b;
    const int mid = b + (e - b) / 2;
    if (e - b >= 2) lds_select<NW>(key, loc, lp, tmp, s_cnt, b, e - 1, mid, up, err);
    return mid;
}

struct FTask {           // a select finished in LDS
    int64_t L, R, nth;   // (small tasks; big ones read L, R from their state)
    int32_t axis, up;
    int32_t big;         // index of the big task's state, or -1
    int32_t pad;
};

// One workgroup per select: keys of [L, R] into LDS, the quickselect, then
// every field of the records permuted into buffer 0. kCap (at most kFin)
// sizes the LDS arrays: a phase whose selects are all known to have at most
// 2048 or 1024 elements runs a smaller form (28 or 14 KB: five or eleven
// workgroups per CU instead of two).
template <int kCap>
__global__ __launch_bounds__(256) void k_lds_select(Bufs bufs, const FTask* __restrict__ ft, uint32_t nft,
                                                    const RState* __restrict__ st_big, uint32_t nbig,
                                                    uint32_t* __restrict__ err) {
    static_assert(kCap <= kFin, "select capacity");
    __shared__ double s_key[kCap];
    __shared__ uint16_t s_loc[kCap], s_lp[kCap], s_tmp[kCap];
    __shared__ uint32_t s_cnt[4];
    const uint32_t ti = blockIdx.x;
    if (ti >= nft) return;
    const FTask T = ft[ti];
    int64_t L = T.L, R = T.R;
    int par = 0;
    if (T.big >= 0) {
        if ((uint32_t)T.big >= nbig) {
            if (threadIdx.x == 0) raise_err(err, kErrLds);
            return;
        }
        const RState S = st_big[T.big];
        if (!S.fin) return;
        L = S.L;
        R = S.R;
        par = S.par;
    }
    const int64_t m = R - L + 1;
    if (m < 2 || m > kCap || L < 0 || R >= bufs.n || T.nth < L || T.nth > R || (par & ~1)) {
        if (threadIdx.x == 0) raise_err(err, kErrLds);
        return;
    }
    const int mi = (int)m;
    const double* key = axis_of(bufs, par, T.axis);
    for (int i = threadIdx.x; i < mi; i += 256) {
        s_key[i] = key[L + i];
        s_loc[i] = (uint16_t)i;
    }
    __syncthreads();
    lds_select<4>(s_key, s_loc, s_lp, s_tmp, s_cnt, 0, mi - 1, (int)(T.nth - L), T.up, err);
    // apply the arrangement to every field (staged through s_key: the source
    // may be buffer 0 itself)
    for (int f = 0; f < 3; ++f) {
        const double* src = axis_of(bufs, par, f);
        double* dst = f == 0 ? bufs.x[0] : (f == 1 ? bufs.y[0] : bufs.z[0]);
        __syncthreads();
        for (int i = threadIdx.x; i < mi; i += 256) s_key[i] = src[L + i];
        __syncthreads();
        for (int i = threadIdx.x; i < mi; i += 256) {
            const int k = s_loc[i];
            if (k < mi) dst[L + i] = s_key[k];
        }
    }
    int32_t* s_id = reinterpret_cast<int32_t*>(s_key);
    __syncthreads();
    const int32_t* src_id = bid(bufs, par);
    for (int i = threadIdx.x; i < mi; i += 256) s_id[i] = src_id[L + i];
    __syncthreads();
    for (int i = threadIdx.x; i < mi; i += 256) {
        const int k = s_loc[i];
        if (k < mi) bufs.id[0][L + i] = s_id[k];
    }
}

// ---------------------------------------------------------------------------
// Mid-size selects (kFin < m <= kWgMax): every round in one workgroup
// ---------------------------------------------------------------------------
// The same Lomuto rounds as the device-wide kernels (flags, ranks, the slots
// of the not-less elements, the moves with their backward chains), each pass
// a loop of the workgroup's 1024 threads over the live range, separated by
// barriers: no launch, state round trip or host check per round. The flag
// words of the range and their prefix counts stay in LDS, so a chain's hops
// read LDS. Once the range holds at most kFin elements the select finishes
// in LDS like k_lds_select, in the same workgroup.
constexpr int kWgThreads = 1024;
constexpr int kWgWaves = kWgThreads / 64;
constexpr int kWgWords = 4096;                         // flag words in LDS
constexpr int64_t kWgMax = (int64_t)kWgWords * 64;    // 262,144 elements

struct WTask {           // a mid-size select of one phase (host-built)
    int64_t b, e, nth;
    int32_t axis, up;
};

struct WgLds {
    union {
        struct {
            uint64_t mask[kWgWords];  // word w: slots L + 64w .. L + 64w + 63, bit set = less
            uint32_t pre[kWgWords];   // exclusive prefix of the words' counts
        } r;
        struct {
            double key[kFin];
            uint16_t loc[kFin], lp[kFin], tmp[kFin];
        } f;
    } u;
    uint32_t wave[kWgWaves];
    uint32_t cnt[kWgWaves];
    int64_t st[4];  // L, R, lim, next action (broadcast by thread 0)
};

// Copies [lo, hi) of every field from buffer `from` to buffer 0, by the workgroup.
__device__ inline void wg_copy_to0(const Bufs& b, int from, int64_t lo, int64_t hi) {
    if (from == 0) return;
    for (int64_t p = lo + threadIdx.x; p < hi; p += kWgThreads) copy_rec(b, from, 0, p, p);
}

__global__ __launch_bounds__(kWgThreads) void k_wg_select(Bufs bufs, const WTask* __restrict__ wt, uint32_t nwt,
                                                          uint32_t* __restrict__ npos, uint32_t* __restrict__ err) {
    __shared__ WgLds S;
    const uint32_t ti = blockIdx.x;
    if (ti >= nwt) return;
    const WTask T = wt[ti];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (T.b < 0 || T.e > bufs.n || T.e - T.b <= kFin || T.e - T.b > kWgMax || T.nth < T.b || T.nth >= T.e ||
        T.axis < 0 || T.axis > 2) {
        if (tid == 0) raise_err(err, kErrRound);
        return;
    }
    int64_t L = T.b, R = T.e - 1;
    int par = 0;
    bool finish = false;  // [L, R] left for the LDS finish
    const int64_t bound = T.e - T.b;
    int64_t round = 0;
    for (; round <= bound; ++round) {
        const double* key = axis_of(bufs, par, T.axis);
        const int64_t pv = L + (R - L) / 2;
        const double pk = key[pv];
        const double kr = key[R];
        const int nw = (int)((R - L + 63) / 64);  // words over [L, R): slot R holds the pivot
        // flags and per-word counts, four words of a wavefront per step so
        // their key loads are in flight together
        for (int w0 = wave; w0 < nw; w0 += 4 * kWgWaves) {
            double kv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t p = L + (int64_t)(w0 + k * kWgWaves) * 64 + lane;
                kv[k] = p < R ? (p == pv ? kr : key[p]) : 0.0;
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int w = w0 + k * kWgWaves;
                const int64_t p = L + (int64_t)w * 64 + lane;
                const bool f = p < R && key_less(kv[k], pk, T.up);
                const uint64_t m = __ballot(f);
                if (lane == 0 && w < nw) {
                    S.u.r.mask[w] = m;
                    S.u.r.pre[w] = (uint32_t)__popcll(m);
                }
            }
        }
        __syncthreads();
        // exclusive prefix of the counts: four words per thread
        {
            constexpr int kPer = kWgWords / kWgThreads;
            uint32_t c[kPer], sum = 0;
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int w = tid * kPer + k;
                c[k] = w < nw ? S.u.r.pre[w] : 0u;
                sum += c[k];
            }
            uint32_t total;
            uint32_t run = wg_exclusive(sum, S.wave, &total);
#pragma unroll
            for (int k = 0; k < kPer; ++k) {
                const int w = tid * kPer + k;
                if (w < nw) S.u.r.pre[w] = run;
                run += c[k];
            }
            if (tid == 0) S.st[2] = L + (int64_t)total;
        }
        __syncthreads();
        const int64_t lim = S.st[2];
        if (lim < L || lim > R) {
            if (tid == 0) raise_err(err, kErrRound);
            return;
        }
        auto rank_of = [&](int64_t q, uint64_t* mword) {
            const int64_t o = q - L;
            const uint64_t m = S.u.r.mask[o >> 6];
            *mword = m;
            return (int64_t)S.u.r.pre[o >> 6] + __popcll(m & ((1ull << (o & 63)) - 1ull));
        };
        // the slots of the not-less elements, in order (offsets from T.b)
        for (int64_t p = L + tid; p < R; p += kWgThreads) {
            uint64_t m;
            const int64_t rk = rank_of(p, &m);
            if ((m >> ((p - L) & 63)) & 1ull) continue;
            const int64_t j = (p - L) - rk;
            if (j >= 0 && L + j < R) npos[L + j] = (uint32_t)(p - T.b);
        }
        __syncthreads();
        // the round's permutation into the other buffer (see k_rd_move): the
        // direct moves first, four slots per thread per step with their
        // loads issued together (a less element to the front, a not-less one
        // at or above lim in place), then the slots whose element is replaced
        // by the end of a backward chain
        const int src = par, dst = par ^ 1;
        const int64_t span = R - L;
        for (int64_t p0 = L + tid; p0 < R; p0 += 4 * kWgThreads) {
            int64_t from[4], to[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t p = p0 + (int64_t)k * kWgThreads;
                from[k] = -1;
                to[k] = -1;
                if (p >= R) continue;
                uint64_t m;
                const int64_t rk = rank_of(p, &m);
                const bool f = (m >> ((p - L) & 63)) & 1ull;
                if (f) {
                    const int64_t d = L + rk;
                    if (d >= L && d < lim) {
                        from[k] = p == pv ? R : p;
                        to[k] = d;
                    } else {
                        raise_err(err, kErrRound);
                    }
                } else if (p >= lim) {
                    from[k] = p == pv ? R : p;  // the slot pv holds what was at R
                    to[k] = p == lim ? R : p;
                }
            }
            double vx[4], vy[4], vz[4];
            int32_t vi[4];
            const double *sx = bx(bufs, src), *sy = by(bufs, src), *sz = bz(bufs, src);
            const int32_t* si = bid(bufs, src);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int64_t q = from[k] < 0 ? L : from[k];
                vx[k] = sx[q];
                vy[k] = sy[q];
                vz[k] = sz[q];
                vi[k] = si[q];
            }
            double *tx = bx(bufs, dst), *ty = by(bufs, dst), *tz = bz(bufs, dst);
            int32_t* tdi = bid(bufs, dst);
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                if (to[k] < 0) continue;
                tx[to[k]] = vx[k];
                ty[to[k]] = vy[k];
                tz[to[k]] = vz[k];
                tdi[to[k]] = vi[k];
            }
        }
        for (int64_t p = lim + tid; p < R; p += kWgThreads) {
            uint64_t m;
            const int64_t rk = rank_of(p, &m);
            const bool f = (m >> ((p - L) & 63)) & 1ull;
            if (!f) continue;  // moved in place above
            int64_t s = p;
            bool less = f;
            int64_t rs = rk;
            int64_t hops = 0;
            bool bad = false;
            while (less) {
                if (++hops > span) {
                    bad = true;
                    break;
                }
                if (hops <= 3) {
                    s = L + rs;
                } else {
                    const int64_t j = (s - L) - rs;  // not-less slots below s
                    if (j < 1 || j > span) {
                        bad = true;
                        break;
                    }
                    const int64_t P = T.b + (int64_t)npos[L + j - 1];
                    if (P < L || P >= s) {
                        bad = true;
                        break;
                    }
                    s -= j * ((s - P + j - 1) / j);
                    if (s == P) break;
                }
                if (s < L || s >= lim) {
                    bad = true;
                    break;
                }
                uint64_t ms;
                rs = rank_of(s, &ms);
                less = (ms >> ((s - L) & 63)) & 1ull;
            }
            if (bad) raise_err(err, hops > span ? kErrLoop : kErrRound);
            else copy_rec(bufs, src, dst, s == pv ? R : s, p == lim ? R : p);
        }
        if (tid == 0) copy_rec(bufs, src, dst, pv, lim);  // the pivot
        __syncthreads();
        // the next range (src/spatial_sort.cpp:65-69); what leaves it is final
        int64_t nL = L, nR = R;
        int action;  // 0: all of [L, R] final; 1: more rounds; 2: the LDS finish
        if (lim == T.nth) {
            action = 0;
        } else {
            if (T.nth < lim) nR = lim - 1;
            else nL = lim + 1;
            const int64_t size = nR - nL + 1;
            action = size > kFin ? 1 : (size >= 2 ? 2 : 0);
        }
        if (action == 0) {
            wg_copy_to0(bufs, dst, L, R + 1);
            L = R + 1;
            par = 0;
            break;
        }
        wg_copy_to0(bufs, dst, L, nL);
        wg_copy_to0(bufs, dst, nR + 1, R + 1);
        L = nL;
        R = nR;
        par = dst;
        __syncthreads();  // every read of this round's flag words is done
        if (action == 2) {
            finish = true;
            break;
        }
    }
    if (round > bound) {
        if (tid == 0) raise_err(err, kErrLoop);
        return;
    }
    if (!finish) return;
    // the LDS finish (as k_lds_select): keys of [L, R] from buffer par, the
    // quickselect on a 16-bit arrangement, then every field permuted into buffer 0
    const int mi = (int)(R - L + 1);
    const double* key = axis_of(bufs, par, T.axis);
    for (int i = tid; i < mi; i += kWgThreads) {
        S.u.f.key[i] = key[L + i];
        S.u.f.loc[i] = (uint16_t)i;
    }
    __syncthreads();
    lds_select<kWgWaves>(S.u.f.key, S.u.f.loc, S.u.f.lp, S.u.f.tmp, S.cnt, 0, mi - 1, (int)(T.nth - L), T.up, err);
    for (int f = 0; f < 3; ++f) {
        const double* fsrc = axis_of(bufs, par, f);
        double* fdst = f == 0 ? bufs.x[0] : (f == 1 ? bufs.y[0] : bufs.z[0]);
        __syncthreads();
        for (int i = tid; i < mi; i += kWgThreads) S.u.f.key[i] = fsrc[L + i];
        __syncthreads();
        for (int i = tid; i < mi; i += kWgThreads) {
            const int k = S.u.f.loc[i];
            if (k < mi) fdst[L + i] = S.u.f.key[k];
        }
    }
    int32_t* s_id = reinterpret_cast<int32_t*>(S.u.f.key);
    __syncthreads();
    const int32_t* isrc = bid(bufs, par);
    for (int i = tid; i < mi; i += kWgThreads) s_id[i] = isrc[L + i];
    __syncthreads();
    for (int i = tid; i < mi; i += kWgThreads) {
        const int k = S.u.f.loc[i];
        if (k < mi) bufs.id[0][L + i] = s_id[k];
    }
}

// ---------------------------------------------------------------------------
// Whole Hilbert subtrees of at most kSub elements, one wavefront each
// ---------------------------------------------------------------------------
constexpr int kSub = 1024;
constexpr int kStack = 64;
constexpr int kSmall = 64;       // nodes of at most this many elements: one lane each
constexpr int kSmallCap = 128;   // > kSub / 9: small nodes have more than 8 elements, disjoint
constexpr int kSubCaps[3] = {320, 640, kSub};  // the forms' capacities, one per leaf class

struct Leaf {
    int64_t lo, hi;  // [lo, hi)
    int32_t x, ux, uy, uz;
};

// hilbert_split of src/spatial_sort.cpp:74-79 on local slots [b, e), by one
// lane alone: the quickselect of :60-71 with the Lomuto partition of :40-51,
// step for step, on loc[] with the keys read through it.
__device__ int lane_split(const double* key, uint16_t* loc, int b, int e, int up, uint32_t* err) {
    if (b >= e) return b;
    const int mid = b + (e - b) / 2;
    int l = b, r = e - 1;
    if (l == r) return mid;
    for (int guard = 0; guard <= e - b; ++guard) {
        const int pv = l + (r - l) / 2;
        uint16_t t = loc[pv];
        loc[pv] = loc[r];
        loc[r] = t;
        const double pk = key[loc[r]];
        int res = l;
        for (int i = l; i < r; ++i) {
            const uint16_t li = loc[i];
            if (key_less(key[li], pk, up)) {
                loc[i] = loc[res];
                loc[res] = li;
                ++res;
            }
        }
        t = loc[r];
        loc[r] = loc[res];
        loc[res] = t;
        if (res == mid) return mid;
        if (mid < res) r = res - 1;
        else l = res + 1;
    }
    raise_err(err, kErrLoop);
    return mid;
}

// A node of the recursion (src/spatial_sort.cpp:95-117) in one word: slots
// [b, e) (11 bits each), the split axis x and the three direction flags.
__device__ inline uint32_t pack_node(int b, int e, int x, int ux, int uy, int uz) {
    return (uint32_t)b | ((uint32_t)e << 11) | ((uint32_t)x << 22) | ((uint32_t)(ux != 0) << 24) |
           ((uint32_t)(uy != 0) << 25) | ((uint32_t)(uz != 0) << 26);
}

// The eight children of a node split at m0..m8 (src/spatial_sort.cpp:109-116).
__device__ inline void child_nodes(int ch[8][6], const int m[9], int x, int y, int z, int ux, int uy, int uz) {
    const int c[8][6] = {{m[0], m[1], z, uz, ux, uy},   {m[1], m[2], y, uy, uz, ux},
                         {m[2], m[3], y, uy, uz, ux},   {m[3], m[4], x, ux, !uy, !uz},
                         {m[4], m[5], x, ux, !uy, !uz}, {m[5], m[6], y, !uy, uz, !ux},
                         {m[6], m[7], y, !uy, uz, !ux}, {m[7], m[8], z, !uz, !ux, uy}};
    for (int i = 0; i < 8; ++i)
        for (int f = 0; f < 6; ++f) ch[i][f] = c[i][f];
}

// Nodes above kSmall elements are split by the whole wavefront (lds_split);
// their children of more than 8 and at most kSmall elements are set aside,
// and once the stack is empty each lane takes its share of them and runs the
// reference's sequential splits on its own node (disjoint slot ranges, so the
// lanes never touch the same slots, and the order in which nodes are
// processed does not matter). A node of at most 64 elements has children of
// at most 8 (64 -> 32 -> 16 -> 8), so its seven splits finish it.
//
// kCap (at most kSub) sizes the LDS arrays. The plan sorts its leaves into
// size classes and each class runs the smallest form that holds it: the
// 320-element form's 12 KB of LDS allow 3-4 wavefronts per SIMD, the
// 640-element one's 21 KB about 2, the 1,024-element one's 33 KB 1 (every
// leaf of a 10^7-point set has at most 268 elements).
template <int kCap>
__global__ __launch_bounds__(64) void k_subtree(Bufs bufs, const Leaf* __restrict__ leaves, uint32_t nleaves,
                                                int32_t* __restrict__ order, uint32_t* __restrict__ err) {
    static_assert(kCap <= kSub && kCap / 9 < kSmallCap, "subtree capacity");
    __shared__ double s_key[3][kCap];
    __shared__ uint16_t s_loc[kCap], s_lp[kCap], s_tmp[kCap];
    __shared__ int32_t s_stack[kStack][6];
    __shared__ uint32_t s_small[kSmallCap];
    __shared__ uint32_t s_cnt[1];
    __shared__ int s_nsmall;
    for (uint32_t li = blockIdx.x; li < nleaves; li += gridDim.x) {
        const Leaf nd = leaves[li];
        const int64_t m = nd.hi - nd.lo;
        if (m < 1 || m > kCap || nd.lo < 0 || nd.hi > bufs.n || nd.x < 0 || nd.x > 2) {
            if (threadIdx.x == 0) raise_err(err, kErrLds);
            continue;
        }
        const int mi = (int)m;
        __syncthreads();
        for (int i = threadIdx.x; i < mi; i += 64) {
            s_key[0][i] = bufs.x[0][nd.lo + i];
            s_key[1][i] = bufs.y[0][nd.lo + i];
            s_key[2][i] = bufs.z[0][nd.lo + i];
            s_loc[i] = (uint16_t)i;
        }
        int sp = 0;  // the same in every lane
        if (threadIdx.x == 0) {
            s_nsmall = 0;
            if (mi > kSmall) {
                s_stack[0][0] = 0;
                s_stack[0][1] = mi;
                s_stack[0][2] = nd.x;
                s_stack[0][3] = nd.ux;
                s_stack[0][4] = nd.uy;
                s_stack[0][5] = nd.uz;
            } else if (mi > 8) {
                s_small[0] = pack_node(0, mi, nd.x, nd.ux, nd.uy, nd.uz);
                s_nsmall = 1;
            }
        }
        if (mi > kSmall) sp = 1;
        __syncthreads();
        for (int guard = 0; sp > 0 && guard <= mi; ++guard) {  // at most one pop per node
            --sp;
            const int b = s_stack[sp][0], e = s_stack[sp][1];
            const int x = s_stack[sp][2], ux = s_stack[sp][3], uy = s_stack[sp][4], uz = s_stack[sp][5];
            __syncthreads();
            if (e - b <= kSmall || b < 0 || e > mi || x < 0 || x > 2) {
                if (threadIdx.x == 0) raise_err(err, kErrLds);  // only nodes above kSmall are stacked
                continue;
            }
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            int mm[9];
            mm[0] = b;
            mm[8] = e;
            mm[4] = lds_split<1>(s_key[x], s_loc, s_lp, s_tmp, s_cnt, mm[0], mm[8], ux, err);
            mm[2] = lds_split<1>(s_key[y], s_loc, s_lp, s_tmp, s_cnt, mm[0], mm[4], uy, err);
            mm[1] = lds_split<1>(s_key[z], s_loc, s_lp, s_tmp, s_cnt, mm[0], mm[2], uz, err);
            mm[3] = lds_split<1>(s_key[z], s_loc, s_lp, s_tmp, s_cnt, mm[2], mm[4], !uz, err);
            mm[6] = lds_split<1>(s_key[y], s_loc, s_lp, s_tmp, s_cnt, mm[4], mm[8], !uy, err);
            mm[5] = lds_split<1>(s_key[z], s_loc, s_lp, s_tmp, s_cnt, mm[4], mm[6], uz, err);
            mm[7] = lds_split<1>(s_key[z], s_loc, s_lp, s_tmp, s_cnt, mm[6], mm[8], !uz, err);
            int ch[8][6];
            child_nodes(ch, mm, x, y, z, ux, uy, uz);
            int add = 0;
            for (int c = 0; c < 8; ++c) add += (ch[c][1] - ch[c][0] > kSmall) ? 1 : 0;
            if (sp + add > kStack) {
                if (threadIdx.x == 0) raise_err(err, kErrLds);
                break;
            }
            if (threadIdx.x == 0) {
                int w = sp;
                for (int c = 0; c < 8; ++c) {
                    const int sz = ch[c][1] - ch[c][0];
                    if (sz <= 8) continue;
                    if (sz <= kSmall) {
                        if (s_nsmall < kSmallCap) s_small[s_nsmall++] = pack_node(ch[c][0], ch[c][1], ch[c][2], ch[c][3],
                                                                                  ch[c][4], ch[c][5]);
                        else raise_err(err, kErrLds);
                        continue;
                    }
                    for (int f = 0; f < 6; ++f) s_stack[w][f] = ch[c][f];
                    ++w;
                }
            }
            sp += add;
            __syncthreads();
        }
        if (sp > 0 && threadIdx.x == 0) raise_err(err, kErrLoop);
        __syncthreads();
        // the small nodes, one lane each
        const int ns = s_nsmall;
        for (int k = threadIdx.x; k < ns && k < kSmallCap; k += 64) {
            const uint32_t w = s_small[k];
            const int b = (int)(w & 2047u), e = (int)((w >> 11) & 2047u), x = (int)((w >> 22) & 3u);
            const int ux = (int)((w >> 24) & 1u), uy = (int)((w >> 25) & 1u), uz = (int)((w >> 26) & 1u);
            if (b < 0 || e > mi || e - b <= 8 || e - b > kSmall || x > 2) {
                raise_err(err, kErrLds);
                continue;
            }
            const int y = (x + 1) % 3, z = (x + 2) % 3;
            const int m4 = lane_split(s_key[x], s_loc, b, e, ux, err);
            const int m2 = lane_split(s_key[y], s_loc, b, m4, uy, err);
            lane_split(s_key[z], s_loc, b, m2, uz, err);
            lane_split(s_key[z], s_loc, m2, m4, !uz, err);
            const int m6 = lane_split(s_key[y], s_loc, m4, e, !uy, err);
            lane_split(s_key[z], s_loc, m4, m6, uz, err);
            lane_split(s_key[z], s_loc, m6, e, !uz, err);
        }
        __syncthreads();
        for (int i = threadIdx.x; i < mi; i += 64) {
            const int k = s_loc[i];
            if (k < mi) order[nd.lo + i] = bufs.id[0][nd.lo + k];
        }
    }
}

__global__ __launch_bounds__(256) void k_write_order(const int32_t* __restrict__ id, uint64_t n,
                                                     int32_t* __restrict__ order) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (uint64_t)gridDim.x * blockDim.x)
        order[i] = id[i];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
int grid_for(uint64_t work, int per_block, int cus) {
    const uint64_t blocks = (work + per_block - 1) / per_block;
    return (int)std::max<uint64_t>(1, std::min<uint64_t>(blocks, (uint64_t)cus * 32));
}

struct Shuffle {
    uint32_t *j = nullptr, *F = nullptr, *G = nullptr;
};

// The universal shuffle tables for steps 1..K-1 (device arrays of K).
Shuffle build_shuffle(uint64_t K, hipStream_t s, Scratch& sc, uint32_t* err) {
    const int cus = device_cu_count();
    const uint32_t last_step = (uint32_t)(K - 1);
    // expected rejections ~ K^2 / 2^33; a generous margin, redone if short
    const double er = (double)K * (double)K / 8589934592.0;
    uint64_t D = K - 1 + (uint64_t)(er + 12.0 * std::sqrt(er + 1.0)) + 256;
    const uint64_t x0 = ((1ull << 16) | 0x330Eull) & kMask48;  // seed 1 (src/rand48.h:30-33)
    for (int attempt = 0; attempt < 8; ++attempt) {
        uint32_t* u = sc.get<uint32_t>(D);
        const uint64_t runs = (D + kDrawRun - 1) / kDrawRun;
        k_draws<<<(unsigned)((runs + 255) / 256), 256, 0, s>>>(u, D, x0);
        D3_LAUNCH_CHECK();
        const uint32_t thr = 0x80000000u - last_step;  // only u >= 2^31 - (K-1) can be rejected
        const uint64_t nblk = (D + kScanTile - 1) / kScanTile;
        uint32_t* blk = sc.get<uint32_t>(nblk);
        uint32_t* meta = sc.get<uint32_t>(4);  // [0] candidates, [1] rejections, [2] steps
        k_cand_count<<<(unsigned)nblk, kScanThreads, 0, s>>>(u, D, thr, blk);
        D3_LAUNCH_CHECK();
        k_scan_one<<<1, 1024, 0, s>>>(blk, (uint32_t)nblk, meta);
        D3_LAUNCH_CHECK();
        uint32_t ncand = 0;
        copy_d2h(&ncand, meta, 4, s);
        const uint32_t cap = ncand + 1;
        uint32_t* cand = sc.get<uint32_t>(cap);
        uint32_t* cand_u = sc.get<uint32_t>(cap);
        uint32_t* rej = sc.get<uint32_t>(cap);
        k_cand_write<<<(unsigned)nblk, kScanThreads, 0, s>>>(u, D, thr, blk, cand, cand_u, cap, err);
        D3_LAUNCH_CHECK();
        k_resolve<<<1, 64, 0, s>>>(cand_u, D, cand, meta, cap, last_step, rej, meta + 1, err);
        D3_LAUNCH_CHECK();
        uint32_t nrej = 0;
        copy_d2h(&nrej, meta + 1, 4, s);
        if (D < (uint64_t)last_step + nrej) {  // too few draws: redo with more
            D = (uint64_t)last_step + nrej + 4096 + D / 64;
            continue;
        }
        Shuffle T;
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
        fill_u32(T.j, 0u, K, s);
        k_steps<<<grid_for(D, 256, cus), 256, 0, s>>>(u, D, rej, meta + 1, cap, last_step, (uint32_t)K, T.j,
                                                      gcount, err);
        D3_LAUNCH_CHECK();
        const uint64_t kb = (K + kScanTile - 1) / kScanTile;
        uint32_t* kblk = sc.get<uint32_t>(kb);
        k_tile_sum<<<(unsigned)kb, kScanThreads, 0, s>>>(gcount, K, kblk);
        D3_LAUNCH_CHECK();
        k_scan_one<<<1, 1024, 0, s>>>(kblk, (uint32_t)kb, meta + 2);
        D3_LAUNCH_CHECK();
        k_tile_scan<<<(unsigned)kb, kScanThreads, 0, s>>>(gcount, K, kblk, gstart);
        D3_LAUNCH_CHECK();
        k_group_scatter<<<grid_for(K, 256, cus), 256, 0, s>>>(T.j, last_step, (uint32_t)K, gstart, gfill, grp,
                                                              err);
        D3_LAUNCH_CHECK();
        k_group_links<<<grid_for(K, 256, cus), 256, 0, s>>>((uint32_t)K, gstart, gcount, grp, T.F, T.G, err);
        D3_LAUNCH_CHECK();
        return T;
    }
    set_error(DEL3D_ERR_INTERNAL, "spatial sort: could not draw enough shuffle steps");
    throw HipError{DEL3D_ERR_INTERNAL};
}

struct Node {
    int64_t lo, hi;
    int x, ux, uy, uz;
};

struct Split {
    int64_t b, e;
    int axis, up;
};

inline int64_t mid_of(int64_t b, int64_t e) { return b >= e ? b : b + (e - b) / 2; }

// One phase of the plan: its big selects (device rounds) and the selects
// finished in LDS, as offsets into the uploaded tables.
struct Phase {
    size_t task0 = 0, ntasks = 0;   // RTask range
    size_t tile0 = 0, ntiles = 0;   // RTile range
    size_t fin0 = 0, nfin = 0;      // FTask range
    size_t wtask0 = 0, nwtasks = 0; // WTask range
    int64_t max_big = 0;            // largest big select
    int64_t max_fin = 0;            // largest select finished in LDS (kFin for a big one's finish)
};

struct Plan {
    std::vector<RTask> tasks;
    std::vector<RTile> tiles;
    std::vector<FTask> fins;
    std::vector<WTask> wtasks;
    std::vector<Leaf> leaves;       // by class: leaves[class_end[c-1] .. class_end[c]) fit kSubCaps[c]
    std::vector<Phase> phases;
    size_t class_end[3] = {0, 0, 0};
};

// A phase's selects of more than kFin elements run device-wide rounds, all
// of them together; when none of them exceeds kWgMax and there are at least
// wg_min of them (enough to give every CU one) they run instead one workgroup
// each (k_wg_select). A phase that has a larger select anyway pays for the
// device-wide rounds' latency, which the smaller selects then share; a
// phase of few selects would leave most of the chip idle in workgroups.
void add_phase(Plan& P, const std::vector<Split>& splits, int64_t wg_min) {
    Phase ph;
    ph.task0 = P.tasks.size();
    ph.tile0 = P.tiles.size();
    ph.fin0 = P.fins.size();
    ph.wtask0 = P.wtasks.size();
    bool wg = true;
    int64_t mid = 0;
    for (const Split& sp : splits) {
        wg = wg && sp.e - sp.b <= kWgMax;
        mid += sp.e - sp.b > kFin ? 1 : 0;
    }
    wg = wg && mid >= wg_min;
    for (const Split& sp : splits) {
        const int64_t m = sp.e - sp.b;
        if (m < 2) continue;  // nth_element returns at once (src/spatial_sort.cpp:62-64)
        const int64_t nth = sp.b + m / 2;
        if (m <= kFin) {
            P.fins.push_back(FTask{sp.b, sp.e - 1, nth, sp.axis, sp.up, -1, 0});
            ph.max_fin = std::max(ph.max_fin, m);
            continue;
        }
        if (wg) {
            P.wtasks.push_back(WTask{sp.b, sp.e, nth, sp.axis, sp.up});
            continue;
        }
        RTask t;
        t.b = sp.b;
        t.e = sp.e;
        t.nth = nth;
        t.axis = sp.axis;
        t.up = sp.up;
        t.tile0 = (int64_t)(P.tiles.size() - ph.tile0);
        const int32_t local = (int32_t)(P.tasks.size() - ph.task0);
        for (int64_t st = sp.b; st < sp.e; st += kTile)
            P.tiles.push_back(RTile{st, std::min<int64_t>(kTile, sp.e - st), local, 0});
        t.ntiles = (int64_t)(P.tiles.size() - ph.tile0) - t.tile0;
        P.tasks.push_back(t);
        P.fins.push_back(FTask{0, 0, nth, sp.axis, sp.up, local, 0});
        ph.max_fin = kFin;  // its remaining range, known on the device only
        ph.max_big = std::max(ph.max_big, m);
    }
    ph.ntasks = P.tasks.size() - ph.task0;
    ph.ntiles = P.tiles.size() - ph.tile0;
    ph.nfin = P.fins.size() - ph.fin0;
    ph.nwtasks = P.wtasks.size() - ph.wtask0;
    if (ph.ntasks || ph.nfin || ph.nwtasks) P.phases.push_back(ph);
}

// The data-independent recursion plan of every set
// (src/spatial_sort.cpp:95-137).
Plan make_plan(const int64_t* off, int64_t n_sets, int64_t wg_min) {
    Plan P;
    std::vector<Node> level;
    auto place = [&](const Node& nd, std::vector<Node>& next) {
        const int64_t m = nd.hi - nd.lo;
        if (m <= 8) return;  // kHilbertLimit: left as it is
        if (m <= kSub) P.leaves.push_back(Leaf{nd.lo, nd.hi, nd.x, nd.ux, nd.uy, nd.uz});
        else next.push_back(nd);
    };
    for (int64_t si = 0; si < n_sets; ++si) {
        int64_t b = off[si], e = off[si + 1];
        if (e - b < 2) continue;  // src/spatial_sort.cpp:142
        for (;;) {  // multiscale segments: [mid, e) sorted, [b, mid) recursed
            int64_t mid = b;
            if (e - b >= 64) mid = b + (int64_t)((double)(e - b) * 0.125);
            place(Node{mid, e, 0, 0, 0, 0}, level);
            if (mid == b) break;
            e = mid;
        }
    }
    while (!level.empty()) {
        std::vector<Split> px, py, pz;
        std::vector<Node> next;
        for (const Node& nd : level) {
            const int x = nd.x, y = (x + 1) % 3, z = (x + 2) % 3;
            const int ux = nd.ux, uy = nd.uy, uz = nd.uz;
            const int64_t m0 = nd.lo, m8 = nd.hi, m4 = mid_of(m0, m8);
            const int64_t m2 = mid_of(m0, m4), m6 = mid_of(m4, m8);
            const int64_t m1 = mid_of(m0, m2), m3 = mid_of(m2, m4);
            const int64_t m5 = mid_of(m4, m6), m7 = mid_of(m6, m8);
            px.push_back(Split{m0, m8, x, ux});
            py.push_back(Split{m0, m4, y, uy});
            py.push_back(Split{m4, m8, y, !uy});
            pz.push_back(Split{m0, m2, z, uz});
            pz.push_back(Split{m2, m4, z, !uz});
            pz.push_back(Split{m4, m6, z, uz});
            pz.push_back(Split{m6, m8, z, !uz});
            place(Node{m0, m1, z, uz, ux, uy}, next);
            place(Node{m1, m2, y, uy, uz, ux}, next);
            place(Node{m2, m3, y, uy, uz, ux}, next);
            place(Node{m3, m4, x, ux, !uy, !uz}, next);
            place(Node{m4, m5, x, ux, !uy, !uz}, next);
            place(Node{m5, m6, y, !uy, uz, !ux}, next);
            place(Node{m6, m7, y, !uy, uz, !ux}, next);
            place(Node{m7, m8, z, !uz, !ux, uy}, next);
        }
        add_phase(P, px, wg_min);
        add_phase(P, py, wg_min);
        add_phase(P, pz, wg_min);
        level.swap(next);
    }
    // leaves by size class (disjoint ranges: their order does not matter)
    auto cls = [](const Leaf& l) {
        const int64_t m = l.hi - l.lo;
        return m <= kSubCaps[0] ? 0 : (m <= kSubCaps[1] ? 1 : 2);
    };
    std::stable_sort(P.leaves.begin(), P.leaves.end(),
                     [&](const Leaf& a, const Leaf& b) { return cls(a) < cls(b); });
    for (int c = 0; c < 3; ++c) {
        size_t e = c ? P.class_end[c - 1] : 0;
        while (e < P.leaves.size() && cls(P.leaves[e]) == c) ++e;
        P.class_end[c] = e;
    }
    return P;
}

template <class T>
size_t align_up_bytes(size_t n) {
    return (n * sizeof(T) + 255) & ~size_t(255);
}

// Sorts every set: order[off[i] .. off[i+1]) holds indices into the set's
// points, which start at xyz + 3*poff[i] (poff == nullptr: poff = off).
void sort_sets(const double* xyz, int64_t n_pts, const int64_t* poff, const int64_t* off, int64_t n_sets,
               int32_t* order, hipStream_t s) {
    const uint64_t total = (uint64_t)off[n_sets];
    if (total == 0) return;
    const int cus = device_cu_count();
    uint64_t K = 2;
    for (int64_t i = 0; i < n_sets; ++i) K = std::max<uint64_t>(K, (uint64_t)(off[i + 1] - off[i]));

    Scratch sc(s);
    uint32_t* err = sc.get<uint32_t>(1);
    fill_u32(err, 0u, 1, s);
    const Shuffle T = build_shuffle(K, s, sc, err);

    // the plan and the set tables, uploaded in one copy
    const Plan P = make_plan(off, n_sets, cus);
    const size_t b_off = 0;
    const size_t b_poff = b_off + align_up_bytes<int64_t>(n_sets + 1);
    const size_t b_tasks = b_poff + align_up_bytes<int64_t>(n_sets + 1);
    const size_t b_tiles = b_tasks + align_up_bytes<RTask>(P.tasks.size());
    const size_t b_fins = b_tiles + align_up_bytes<RTile>(P.tiles.size());
    const size_t b_wtasks = b_fins + align_up_bytes<FTask>(P.fins.size());
    const size_t b_leaves = b_wtasks + align_up_bytes<WTask>(P.wtasks.size());
    const size_t b_end = b_leaves + align_up_bytes<Leaf>(P.leaves.size());
    std::vector<char> blob(b_end, 0);
    std::memcpy(blob.data() + b_off, off, (n_sets + 1) * sizeof(int64_t));
    std::memcpy(blob.data() + b_poff, poff ? poff : off, (n_sets + 1) * sizeof(int64_t));
    if (!P.tasks.empty()) std::memcpy(blob.data() + b_tasks, P.tasks.data(), P.tasks.size() * sizeof(RTask));
    if (!P.tiles.empty()) std::memcpy(blob.data() + b_tiles, P.tiles.data(), P.tiles.size() * sizeof(RTile));
    if (!P.fins.empty()) std::memcpy(blob.data() + b_fins, P.fins.data(), P.fins.size() * sizeof(FTask));
    if (!P.wtasks.empty())
        std::memcpy(blob.data() + b_wtasks, P.wtasks.data(), P.wtasks.size() * sizeof(WTask));
    if (!P.leaves.empty())
        std::memcpy(blob.data() + b_leaves, P.leaves.data(), P.leaves.size() * sizeof(Leaf));
    char* dblob = sc.get<char>(b_end);
    copy_h2d(dblob, blob.data(), b_end, s);
    const int64_t* doff = reinterpret_cast<const int64_t*>(dblob + b_off);
    const int64_t* dpoff = reinterpret_cast<const int64_t*>(dblob + b_poff);
    const RTask* dtasks = reinterpret_cast<const RTask*>(dblob + b_tasks);
    const RTile* dtiles = reinterpret_cast<const RTile*>(dblob + b_tiles);
    const FTask* dfins = reinterpret_cast<const FTask*>(dblob + b_fins);
    const WTask* dwtasks = reinterpret_cast<const WTask*>(dblob + b_wtasks);
    const Leaf* dleaves = reinterpret_cast<const Leaf*>(dblob + b_leaves);

    Bufs bufs;
    for (int i = 0; i < 2; ++i) {
        bufs.x[i] = sc.get<double>(total);
        bufs.y[i] = sc.get<double>(total);
        bufs.z[i] = sc.get<double>(total);
        bufs.id[i] = sc.get<int32_t>(total);
    }
    bufs.n = (int64_t)total;
    k_place<<<grid_for(total, 256, cus), 256, 0, s>>>(doff, dpoff, (uint32_t)n_sets, total, n_pts, T.j, T.F,
                                                      T.G, (uint32_t)K, order, xyz, bufs, err);
    D3_LAUNCH_CHECK();

    // round buffers, sized for the largest phase
    size_t max_tasks = 0, max_tiles = 0;
    for (const Phase& ph : P.phases) {
        max_tasks = std::max(max_tasks, ph.ntasks);
        max_tiles = std::max(max_tiles, ph.ntiles);
    }
    RState* st[2] = {nullptr, nullptr};
    uint64_t* mask = nullptr;
    uint16_t* wloc = nullptr;
    uint32_t *tcount = nullptr, *toff = nullptr, *ring = nullptr, *npos = nullptr;
    uint32_t* alist[3] = {nullptr, nullptr, nullptr};  // the tiles of rounds r, r-1 and r-2 (r mod 3)
    uint32_t* acnt = nullptr;
    if (!P.wtasks.empty()) npos = sc.get<uint32_t>(total);
    if (max_tasks) {
        st[0] = sc.get<RState>(max_tasks);
        st[1] = sc.get<RState>(max_tasks);
        mask = sc.get<uint64_t>(max_tiles * kTileWords);
        tcount = sc.get<uint32_t>(max_tiles);
        toff = sc.get<uint32_t>(max_tiles);
        wloc = sc.get<uint16_t>(max_tiles * kTileWords);
        if (!npos) npos = sc.get<uint32_t>(total);
        ring = sc.get<uint32_t>(kRing);
        fill_u32(ring, 0u, kRing, s);
        for (auto& l : alist) l = sc.get<uint32_t>(max_tiles);
        acnt = sc.get<uint32_t>(3);
    }
    for (const Phase& ph : P.phases) {
        const RState* final_state = st[0];
        if (ph.nwtasks) {
            k_wg_select<<<(unsigned)ph.nwtasks, kWgThreads, 0, s>>>(bufs, dwtasks + ph.wtask0,
                                                                   (uint32_t)ph.nwtasks, npos, err);
            D3_LAUNCH_CHECK();
        }
        if (ph.ntasks) {
            const RTask* pt = dtasks + ph.task0;
            const RTile* ptl = dtiles + ph.tile0;
            const uint32_t nt = (uint32_t)ph.ntasks, ntl = (uint32_t)ph.ntiles;
            k_rd_init<<<grid_for(nt, 256, cus), 256, 0, s>>>(pt, nt, st[0]);
            D3_LAUNCH_CHECK();
            fill_u32(acnt, 0u, 3, s);
            // Each round works on the tiles that intersect a live range
            // (k_rd_list): its count over the previous round's tiles (which
            // hold the margins that round finalized), its moves over its own;
            // fixed grids stride over the lists, so a round's cost follows the
            // live ranges, not the phase's original size.
            const unsigned g_count = (unsigned)std::min<uint64_t>(ntl, (uint64_t)cus * 8);
            const unsigned g_move = (unsigned)std::min<uint64_t>((uint64_t)ntl * kSlices, (uint64_t)cus * 64);
            const unsigned g_list = grid_for(ntl, 256, cus);
            // A select's range shrinks by at least one slot per round; the
            // expected count is ~1.5 log2(m / kFin) rounds. Check every two
            // rounds once that many have run.
            const int64_t bound = ph.max_big + 2;
            int64_t check_at = 2;
            while ((int64_t(kFin) << check_at) < ph.max_big && check_at < 40) ++check_at;
            check_at = check_at + check_at / 2;
            uint32_t round = 0;
            for (;;) {
                const int c = round & 1;
                const int lr = (int)(round % 3), lp = round ? (int)((round + 2) % 3) : lr;
                k_rd_list<<<g_list, 256, 0, s>>>(ptl, ntl, nt, st[c], alist[lr], acnt + lr);
                D3_LAUNCH_CHECK();
                k_rd_count<<<g_count, kTileThreads, 0, s>>>(bufs, pt, nt, ptl, ntl, st[c], alist[lp], acnt + lp,
                                                            acnt + (round + 1) % 3, mask, wloc, tcount, ring, round,
                                                            err);
                D3_LAUNCH_CHECK();
                k_rd_scan<<<nt, 256, 0, s>>>(pt, nt, ntl, st[c], st[c ^ 1], tcount, toff, err);
                D3_LAUNCH_CHECK();
                k_rd_npos<<<g_move, kTileThreads, 0, s>>>(pt, nt, ptl, ntl, st[c], alist[lr], acnt + lr, mask, wloc,
                                                          toff, npos, bufs.n);
                D3_LAUNCH_CHECK();
                k_rd_move<<<g_move, kTileThreads, 0, s>>>(bufs, pt, nt, ptl, ntl, st[c], st[c ^ 1], alist[lr],
                                                          acnt + lr, mask, wloc, toff, npos, ring, round, err);
                D3_LAUNCH_CHECK();
                ++round;
                if ((int64_t)round >= check_at) {
                    uint32_t active = 0;
                    copy_d2h(&active, ring + (round % kRing), 4, s);
                    if (active == 0) break;
                    check_at = round + 2;
                }
                if ((int64_t)round > bound) {
                    set_error(DEL3D_ERR_INTERNAL, "spatial sort: a select did not finish in %lld rounds",
                              (long long)bound);
                    throw HipError{DEL3D_ERR_INTERNAL};
                }
            }
            // the margins of the last round, over its tiles
            const int last = (int)((round + 2) % 3);
            k_rd_count<<<g_count, kTileThreads, 0, s>>>(bufs, pt, nt, ptl, ntl, st[round & 1], alist[last],
                                                        acnt + last, nullptr, mask, wloc, tcount, ring, round, err);
            D3_LAUNCH_CHECK();
            final_state = st[round & 1];
        }
        if (ph.nfin) {
            if (ph.max_fin <= 1024)
                k_lds_select<1024><<<(unsigned)ph.nfin, 256, 0, s>>>(bufs, dfins + ph.fin0, (uint32_t)ph.nfin,
                                                                     final_state, (uint32_t)ph.ntasks, err);
            else if (ph.max_fin <= 2048)
                k_lds_select<2048><<<(unsigned)ph.nfin, 256, 0, s>>>(bufs, dfins + ph.fin0, (uint32_t)ph.nfin,
                                                                     final_state, (uint32_t)ph.ntasks, err);
            else
                k_lds_select<kFin><<<(unsigned)ph.nfin, 256, 0, s>>>(bufs, dfins + ph.fin0, (uint32_t)ph.nfin,
                                                                     final_state, (uint32_t)ph.ntasks, err);
            D3_LAUNCH_CHECK();
        }
    }
    k_write_order<<<grid_for(total, 256, cus), 256, 0, s>>>(bufs.id[0], total, order);
    D3_LAUNCH_CHECK();
    for (int c = 0; c < 3; ++c) {
        const size_t l0 = c ? P.class_end[c - 1] : 0, nl = P.class_end[c] - l0;
        if (!nl) continue;
        const unsigned g = (unsigned)std::min<size_t>(nl, (size_t)cus * 16);
        if (c == 0) k_subtree<kSubCaps[0]><<<g, 64, 0, s>>>(bufs, dleaves + l0, (uint32_t)nl, order, err);
        else if (c == 1) k_subtree<kSubCaps[1]><<<g, 64, 0, s>>>(bufs, dleaves + l0, (uint32_t)nl, order, err);
        else k_subtree<kSubCaps[2]><<<g, 64, 0, s>>>(bufs, dleaves + l0, (uint32_t)nl, order, err);
        D3_LAUNCH_CHECK();
    }
    uint32_t herr = 0;
    copy_d2h(&herr, err, 4, s);  // also the final wait: the scratch is released after this
    if (herr & kErrIndex) {
        set_error(DEL3D_ERR_INVALID, "spatial sort: an order entry lies outside its set's points");
        throw HipError{DEL3D_ERR_INVALID};
    }
    if (herr) {
        set_error(DEL3D_ERR_INTERNAL, "spatial sort: internal check failed (code %u)", herr);
        throw HipError{DEL3D_ERR_INTERNAL};
    }
}

}  // namespace

void spatial_sort_sets_device(const double* xyz, const int64_t* off, int64_t n_sets, int32_t* order,
                              hipStream_t s) {
    sort_sets(xyz, off[n_sets], nullptr, off, n_sets, order, s);
}

// The reference's (points, order) pair: n_order indices into n_pts points.
void spatial_sort_device(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order, hipStream_t s) {
    if (n_order < 2) return;
    const int64_t off[2] = {0, n_order};
    const int64_t poff[2] = {0, n_pts};
    sort_sets(xyz, n_pts, poff, off, 1, order, s);
}

}  // namespace d3
