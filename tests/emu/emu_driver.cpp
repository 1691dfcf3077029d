// Entry points of the emulated build (test infrastructure): the few host
// helpers the kernels' host code needs, and C wrappers for ctypes.
#include <cstdint>
#include <string>

#include "hipemu.h"

namespace d3 {
std::string& last_error_slot() {
    static thread_local std::string s;
    return s;
}
int device_cu_count() { return emu::g_cus; }
void spatial_sort_device(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                         hipStream_t s);
void spatial_sort_sets_device(const double* xyz, const int64_t* off, int64_t n_sets,
                              int32_t* order, hipStream_t s);
}  // namespace d3

extern "C" int emu_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order) {
    try {
        d3::spatial_sort_device(xyz, n_pts, order, n_order, nullptr);
        return 0;
    } catch (...) {
        return 1;
    }
}

extern "C" int emu_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                                     int32_t* order) {
    try {
        d3::spatial_sort_sets_device(xyz, off, n_sets, order, nullptr);
        return 0;
    } catch (...) {
        return 1;
    }
}
