// trace_recorder.cpp - records every orient3d / in_sphere call the reference
// Delaunay driver makes, as (op, point indices, returned sign).
//
// TEST INFRASTRUCTURE ONLY. Linked by oracle/Makefile against the reference's
// unchanged objects with
//   -Wl,--wrap=_ZN5del3d8orient3dEPKdS1_S1_S1_
//   -Wl,--wrap=_ZN5del3d9in_sphereEPKdS1_S1_S1_S1_
// so the driver's calls (src/triangulation.cpp:430,476,596,600,793) pass
// through the wrappers below; calls internal to src/predicates.cpp are
// resolved inside that object and are not recorded. Every predicate argument
// is a pointer into the caller's point vector (src/triangulation.cpp:136-141),
// so (ptr - base) / 3 recovers the point index exactly.
//
// Usage: trace_recorder N SEED OUT
//   Points: N points whose coordinates are consecutive rand48 states (seeded
//   like srand48(SEED)) times 2^-48, consumed x, y, z per point.
//   OUT: "D3TR" u32 version=1, i64 n_points, i64 n_records, then the points
//   (3*N f64), op[n_records] (u8: 0 orient3d, 1 in_sphere),
//   sign[n_records] (i8), idx[5*n_records] (i32, orient3d leaves idx[4]=-1).
#include "del3d/delaunay.h"
#include "del3d/predicates.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

static const double* g_base = nullptr;
static std::vector<uint8_t> g_op;
static std::vector<int8_t> g_sign;
static std::vector<int32_t> g_idx;

static int32_t index_of(const double* p) { return static_cast<int32_t>((p - g_base) / 3); }

extern "C" {
del3d::Sign __real__ZN5del3d8orient3dEPKdS1_S1_S1_(const double*, const double*, const double*,
                                                   const double*);
del3d::Sign __real__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(const double*, const double*,
                                                       const double*, const double*,
                                                       const double*);

del3d::Sign __wrap__ZN5del3d8orient3dEPKdS1_S1_S1_(const double* a, const double* b,
                                                   const double* c, const double* d) {
    const del3d::Sign s = __real__ZN5del3d8orient3dEPKdS1_S1_S1_(a, b, c, d);
    g_op.push_back(0);
    g_sign.push_back(static_cast<int8_t>(s));
    g_idx.insert(g_idx.end(), {index_of(a), index_of(b), index_of(c), index_of(d), -1});
    return s;
}

del3d::Sign __wrap__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(const double* a, const double* b,
                                                       const double* c, const double* d,
                                                       const double* e) {
    const del3d::Sign s = __real__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(a, b, c, d, e);
    g_op.push_back(1);
    g_sign.push_back(static_cast<int8_t>(s));
    g_idx.insert(g_idx.end(), {index_of(a), index_of(b), index_of(c), index_of(d), index_of(e)});
    return s;
}
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s N SEED OUT\n", argv[0]);
        return 2;
    }
    const int64_t n = std::atoll(argv[1]);
    const uint32_t seed = static_cast<uint32_t>(std::strtoul(argv[2], nullptr, 10));
    std::vector<del3d::Point3> pts(static_cast<size_t>(n));
    uint64_t x = ((static_cast<uint64_t>(seed) << 16) | 0x330Eull) & ((1ull << 48) - 1);
    auto next = [&x] {
        x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
        return static_cast<double>(x) * (1.0 / 281474976710656.0);
    };
    for (auto& p : pts) {
        p.x = next();
        p.y = next();
        p.z = next();
    }
    g_base = &pts[0].x;
    del3d::Delaunay d;
    d.build(pts);

    FILE* f = std::fopen(argv[3], "wb");
    if (!f) return 1;
    const uint32_t version = 1;
    const int64_t nrec = static_cast<int64_t>(g_op.size());
    std::fwrite("D3TR", 1, 4, f);
    std::fwrite(&version, 4, 1, f);
    std::fwrite(&n, 8, 1, f);
    std::fwrite(&nrec, 8, 1, f);
    std::fwrite(pts.data(), sizeof(del3d::Point3), pts.size(), f);
    std::fwrite(g_op.data(), 1, g_op.size(), f);
    std::fwrite(g_sign.data(), 1, g_sign.size(), f);
    std::fwrite(g_idx.data(), 4, g_idx.size(), f);
    std::fclose(f);
    std::printf("points %lld records %lld cells %zu\n", static_cast<long long>(n),
                static_cast<long long>(nrec), d.finite_cells().size());
    return 0;
}
