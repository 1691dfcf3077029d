// delaunay_timing.cpp - times the reference's unchanged Delaunay driver
// (src/triangulation.cpp, Delaunay::build, include/del3d/delaunay.h:62-92)
// on n rand48 points and prints the build time and an FNV-1a hash of the
// finite cells in allocation order. TEST INFRASTRUCTURE: oracle/Makefile
// links it twice, against the reference's own predicates and spatial sort
// (oracle/_ref/ref_delaunay_timing) and against libdel3d_amd.so in their
// place (oracle/_ref/dropin_delaunay_timing), so the two builds can be
// compared for the same cells and for speed.
//
// Points: the rand48 generator seeded 1 (x <- 0x5DEECE66D x + 0xB mod 2^48,
// initial state (1 << 16) | 0x330E), each coordinate = state * 2^-48 after
// one step, x then y then z per point.
#include "del3d/delaunay.h"

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

int main(int argc, char** argv) {
    const long n = argc > 1 ? std::atol(argv[1]) : 10000;
    const int reps = argc > 2 ? std::atoi(argv[2]) : 3;
    std::vector<del3d::Point3> pts((size_t)n);
    uint64_t x = (1ull << 16) | 0x330Eull;
    auto next = [&x] {
        x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
        return (double)x * 0x1p-48;
    };
    for (auto& p : pts) {
        p.x = next();
        p.y = next();
        p.z = next();
    }
    double best = 1e30;
    uint64_t h = 0;
    size_t cells = 0;
    // one untimed build first: it pays one-time set-up (for the GPU library,
    // the runtime's initialisation on the first spatial sort)
    for (int r = -1; r < reps; ++r) {
        del3d::Delaunay d;
        const auto t0 = std::chrono::steady_clock::now();
        d.build(pts);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (r >= 0 && s < best) best = s;
        h = 1469598103934665603ull;
        for (const auto& t : d.finite_cells())
            for (int v : t)
                for (int b = 0; b < 4; ++b) {
                    h ^= ((uint64_t)(uint32_t)v >> (8 * b)) & 0xffull;
                    h *= 1099511628211ull;
                }
        cells = d.finite_cells().size();
    }
    std::printf("{\"points\": %ld, \"cells\": %zu, \"hash\": \"%016llx\", \"build_s\": %.6f, \"reps\": %d}\n", n,
                cells, (unsigned long long)h, best, reps);
    return 0;
}
