// calib.hip - measurement kernels, built into their own library
// (libdel3d_calib.so), never part of the product: a plain device-to-device
// copy of known size, which calibrates the HBM byte counters (FETCH_SIZE /
// WRITE_SIZE) that profile the product kernels, and a gather-only pass over a
// predicate batch, which measures the ceiling random point gathers set.
#include <hip/hip_runtime.h>

#include <cstdint>

__global__ __launch_bounds__(256) void calib_copy(const uint4* __restrict__ src,
                                                  uint4* __restrict__ dst, size_t n16) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
         i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

extern "C" int del3d_calib_copy(const void* src, void* dst, size_t bytes, void* stream) {
    const size_t n16 = bytes / 16;
    calib_copy<<<2048, 256, 0, (hipStream_t)stream>>>(static_cast<const uint4*>(src),
                                                     static_cast<uint4*>(dst), n16);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}

// The point gathers of a predicate batch and nothing else: each lane reads
// its query's index record and op byte and loads the query's 4 (orient3d) or
// 5 points the way the predicate kernel does (a 16- and an 8-byte load per
// point), summing the coordinates so the loads stay live. Its time is the
// gather-only ceiling of the same batch on this chip. Indices outside
// [0, n_pts) are skipped.
__global__ __launch_bounds__(256) void calib_gather(const double* __restrict__ xyz, uint32_t n_pts,
                                                    const uint8_t* __restrict__ op,
                                                    const int32_t* __restrict__ idx, int64_t n_q,
                                                    double* __restrict__ sink) {
    double acc = 0.0;
    for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n_q;
         q += (int64_t)gridDim.x * blockDim.x) {
        const int np = op[q] == 0 ? 4 : 5;
#pragma unroll
        for (int k = 0; k < 5; ++k) {
            const uint32_t i = (uint32_t)idx[5 * q + k];
            if (k < np && i < n_pts) {
                const double2 a = *reinterpret_cast<const double2*>(xyz + 3 * (size_t)i);
                acc += a.x + a.y + xyz[3 * (size_t)i + 2];
            }
        }
    }
    if (acc == 1.0e300) sink[0] = acc;  // never true for finite points; keeps the loads
}

extern "C" int del3d_calib_gather(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                                  int64_t n_q, double* sink, void* stream) {
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, 0) != hipSuccess) return 1;
    const int64_t want = (n_q + 255) / 256;
    const int blocks = (int)(want < (int64_t)cu * 16 ? want : (int64_t)cu * 16);
    if (blocks <= 0) return 0;
    calib_gather<<<blocks, 256, 0, (hipStream_t)stream>>>(xyz, (uint32_t)n_pts, op, idx, n_q, sink);
    return hipGetLastError() == hipSuccess ? 0 : 1;
}
