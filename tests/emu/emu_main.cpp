// Emulated (CPU, AddressSanitizer) run of the GPU spatial sort, checked
// against the oracle (test infrastructure). Usage: emu_sort N [N ...]
// Environment: EMU_DEFER=1 (queued kernels, immediate pageable copies and
// frees), EMU_BLOCK_ORDER=0|1|2.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../oracle/oracle.h"

extern "C" int emu_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order);
extern "C" int emu_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                                     int32_t* order);

int main(int argc, char** argv) {
    int fails = 0;
    for (int a = 1; a < argc; ++a) {
        const int64_t n = std::atoll(argv[a]);
        std::vector<double> xyz(3 * (size_t)n + 3);
        orc_pts_random_cube(n, 7u, xyz.data());
        // every fourth point snapped to a coarse grid: ties on every axis
        for (int64_t i = 0; i < n; i += 4)
            for (int k = 0; k < 3; ++k) xyz[3 * i + k] = (double)(int)(xyz[3 * i + k] * 4) / 4;
        std::vector<int32_t> got(n), want(n);
        for (int64_t i = 0; i < n; ++i) got[i] = want[i] = (int32_t)i;
        if (emu_spatial_sort(xyz.data(), n, got.data(), n) != 0) { std::printf("n=%lld error\n", (long long)n); ++fails; continue; }
        orc_spatial_sort(xyz.data(), n, want.data());
        const bool ok = got == want;
        std::printf("n=%lld %s\n", (long long)n, ok ? "ok" : "MISMATCH");
        fails += ok ? 0 : 1;
    }
    // three sets in one call
    const int64_t sizes[3] = {1500, 3, 700};
    std::vector<double> xyz;
    std::vector<int64_t> off = {0};
    std::vector<int32_t> got, want;
    for (int s = 0; s < 3; ++s) {
        std::vector<double> p(3 * (size_t)sizes[s] + 3);
        orc_pts_random_cube(sizes[s], 30u + s, p.data());
        xyz.insert(xyz.end(), p.begin(), p.begin() + 3 * sizes[s]);
        std::vector<int32_t> o(sizes[s]);
        for (int64_t i = 0; i < sizes[s]; ++i) o[i] = (int32_t)i;
        got.insert(got.end(), o.begin(), o.end());
        orc_spatial_sort(p.data(), sizes[s], o.data());
        want.insert(want.end(), o.begin(), o.end());
        off.push_back(off.back() + sizes[s]);
    }
    const bool sets_ok = emu_spatial_sort_sets(xyz.data(), off.data(), 3, got.data()) == 0 && got == want;
    std::printf("sets %s\n", sets_ok ? "ok" : "MISMATCH");
    fails += sets_ok ? 0 : 1;
    return fails ? 1 : 0;
}
