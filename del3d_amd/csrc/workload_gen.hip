// workload_gen.hip - synthetic query streams for the benchmarks, generated on
// the device (libdel3d_workloads.so; not part of the product library).
//
// Query q of a batch draws its op code and five set-local point indices from
// a counter-based hash, so any query can be regenerated on the host with the
// same integer arithmetic (del3d_amd/workloads.py: hashed_queries):
//   h(q, k) = splitmix64(seed * 2^40 + 8*q + k), k = 0..5
//   idx[q][k] = (hi32(h(q, k)) * n_set) >> 32       k = 0..4
//   op[q]     = h(q, 5) & 1                          (0 orient3d, 1 in_sphere)
// where n_set is the point count of the set holding q (sets by q_off).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

__device__ inline uint64_t splitmix64(uint64_t x) {
    uint64_t z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void gen_queries(uint64_t seed, const int64_t* __restrict__ pt_off,
                                                   const int64_t* __restrict__ q_off, uint32_t n_sets,
                                                   uint64_t n_q, uint8_t* __restrict__ op,
                                                   int32_t* __restrict__ idx) {
    for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_q;
         q += (uint64_t)gridDim.x * blockDim.x) {
        uint32_t lo = 0, hi = n_sets;  // last s with q_off[s] <= q
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) >> 1;
            if ((uint64_t)q_off[mid] <= q) lo = mid;
            else hi = mid;
        }
        const uint64_t n = (uint64_t)(pt_off[lo + 1] - pt_off[lo]);
        const uint64_t base = seed * (1ull << 40) + 8 * q;
#pragma unroll
        for (int k = 0; k < 5; ++k) idx[5 * q + k] = (int32_t)(((splitmix64(base + k) >> 32) * n) >> 32);
        op[q] = (uint8_t)(splitmix64(base + 5) & 1u);
    }
}

}  // namespace

// pt_off, q_off: device arrays of n_sets+1; op, idx: device outputs.
extern "C" int del3d_gen_queries(uint64_t seed, const int64_t* pt_off, const int64_t* q_off,
                                 int64_t n_sets, int64_t n_q, uint8_t* op, int32_t* idx,
                                 void* stream) {
    if (n_q <= 0) return 0;
    gen_queries<<<4096, 256, 0, (hipStream_t)stream>>>(seed, pt_off, q_off, (uint32_t)n_sets,
                                                        (uint64_t)n_q, op, idx);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
