// calib.hip - a plain device-to-device copy kernel of known size, built into
// its own library (libdel3d_calib.so), used only to calibrate the HBM byte
// counters (FETCH_SIZE / WRITE_SIZE) that profile the product kernels.
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
