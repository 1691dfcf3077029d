// Entry points of the emulated build (test infrastructure): C wrappers over
// the library's C ABI, run on the CPU by hipemu.h.
#include <cstdint>
#include <cstdlib>

#include "../../include/del3d_amd.h"

extern "C" int emu_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order) {
    setenv("DEL3D_ENABLE_GPU_SORT", "1", 1);
    return del3d_spatial_sort_host(xyz, n_pts, order, n_order);
}

extern "C" int emu_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                                     int32_t* order) {
    setenv("DEL3D_ENABLE_GPU_SORT", "1", 1);
    // device buffers are host memory in the emulation
    return del3d_spatial_sort_sets(xyz, off, n_sets, order, nullptr);
}

extern "C" int emu_predicates(const double* xyz, int64_t n_pts, const uint8_t* op,
                              const int32_t* idx, int64_t n_q, int8_t* out) {
    return del3d_predicates_host(xyz, n_pts, op, idx, n_q, out);
}

extern "C" int emu_predicates_stats(const double* xyz, int64_t n_pts, const uint8_t* op,
                                    const int32_t* idx, int64_t n_q, int8_t* out, int64_t* st) {
    return del3d_predicates_stats(xyz, n_pts, op, idx, n_q, out, nullptr, st);
}
