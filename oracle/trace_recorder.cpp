// trace_recorder.cpp - records every exact predicate call the reference
// Delaunay driver makes, as (op, point indices, returned result).
//
// TEST INFRASTRUCTURE ONLY. Linked by oracle/Makefile against the reference's
// unchanged objects with -Wl,--wrap= for orient3d, in_sphere and the coplanar
// predicates (include/del3d/predicates.h:26,49-50;
// include/del3d/predicates2.h:33-69), so the driver's calls
// (src/triangulation.cpp:430,476,499-503,528-554,596-624,786,793) pass
// through the wrappers below; calls internal to src/predicates.cpp and
// src/predicates2.cpp are resolved inside those objects and are not recorded.
// Every predicate argument is a pointer into the caller's point vector
// (src/triangulation.cpp:136-141), so (ptr - base) / 3 recovers the point
// index exactly; the recorder aborts if one does not.
//
// Usage: trace_recorder N SEED OUT          N rand48 points (coordinates are
//                                           consecutive states of srand48(SEED)
//                                           times 2^-48, x, y, z per point)
//        trace_recorder lattice NX NY NZ OUT  the NX x NY x NZ integer lattice,
//                                           index (i*NY + j)*NZ + k
// OUT: "D3TR" u32 version=2, i64 n_points, i64 n_records, then the points
// (3*N f64), op[n_records] (u8, the mixed-batch codes of include/del3d_amd.h:
// 0 orient3d, 1 in_sphere, 16 + DEL3D_OP2_* for the coplanar predicates),
// result[n_records] (i8), idx[5*n_records] (i32, unused entries -1).
#include "del3d/delaunay.h"
#include "del3d/predicates.h"
#include "del3d/predicates2.h"

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <vector>

static const double* g_base = nullptr;
static int64_t g_n = 0;
static std::vector<uint8_t> g_op;
static std::vector<int8_t> g_res;
static std::vector<int32_t> g_idx;

static int32_t index_of(const double* p) {
    const int64_t d = p - g_base;
    if (d < 0 || d % 3 != 0 || d / 3 >= g_n) {
        std::fprintf(stderr, "predicate argument is not a point of the input\n");
        std::abort();
    }
    return static_cast<int32_t>(d / 3);
}

static void record(uint8_t op, int res, std::initializer_list<const double*> pts) {
    g_op.push_back(op);
    g_res.push_back(static_cast<int8_t>(res));
    int k = 0;
    for (const double* p : pts) {
        g_idx.push_back(index_of(p));
        ++k;
    }
    for (; k < 5; ++k) g_idx.push_back(-1);
}

using del3d::Sign;
using D = const double*;

// mixed-batch op codes (include/del3d_amd.h)
enum : uint8_t {
    kOrient3d = 0, kInSphere = 1, kBase2 = 16,
    kCopOrient = kBase2 + 0, kCopOrient4 = kBase2 + 1, kCollinear = kBase2 + 2,
    kEqual = kBase2 + 3, kPosition = kBase2 + 4, kInCircle = kBase2 + 5,
    kInCircleRaw = kBase2 + 6,
};

extern "C" {
Sign __real__ZN5del3d8orient3dEPKdS1_S1_S1_(D, D, D, D);
Sign __real__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(D, D, D, D, D);
Sign __real__ZN5del3d20coplanar_orientationEPKdS1_S1_(D, D, D);
Sign __real__ZN5del3d20coplanar_orientationEPKdS1_S1_S1_(D, D, D, D);
bool __real__ZN5del3d9collinearEPKdS1_S1_(D, D, D);
bool __real__ZN5del3d5equalEPKdS1_(D, D);
del3d::CollinearPosition __real__ZN5del3d18collinear_positionEPKdS1_S1_(D, D, D);
Sign __real__ZN5del3d18coplanar_in_circleEPKdS1_S1_S1_(D, D, D, D);
Sign __real__ZN5del3d30coplanar_in_circle_unperturbedEPKdS1_S1_S1_(D, D, D, D);

Sign __wrap__ZN5del3d8orient3dEPKdS1_S1_S1_(D a, D b, D c, D d) {
    const Sign s = __real__ZN5del3d8orient3dEPKdS1_S1_S1_(a, b, c, d);
    record(kOrient3d, s, {a, b, c, d});
    return s;
}
Sign __wrap__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(D a, D b, D c, D d, D e) {
    const Sign s = __real__ZN5del3d9in_sphereEPKdS1_S1_S1_S1_(a, b, c, d, e);
    record(kInSphere, s, {a, b, c, d, e});
    return s;
}
Sign __wrap__ZN5del3d20coplanar_orientationEPKdS1_S1_(D p, D q, D r) {
    const Sign s = __real__ZN5del3d20coplanar_orientationEPKdS1_S1_(p, q, r);
    record(kCopOrient, s, {p, q, r});
    return s;
}
Sign __wrap__ZN5del3d20coplanar_orientationEPKdS1_S1_S1_(D p, D q, D r, D t) {
    const Sign s = __real__ZN5del3d20coplanar_orientationEPKdS1_S1_S1_(p, q, r, t);
    record(kCopOrient4, s, {p, q, r, t});
    return s;
}
bool __wrap__ZN5del3d9collinearEPKdS1_S1_(D p, D q, D r) {
    const bool s = __real__ZN5del3d9collinearEPKdS1_S1_(p, q, r);
    record(kCollinear, s ? 1 : 0, {p, q, r});
    return s;
}
bool __wrap__ZN5del3d5equalEPKdS1_(D p, D q) {
    const bool s = __real__ZN5del3d5equalEPKdS1_(p, q);
    record(kEqual, s ? 1 : 0, {p, q});
    return s;
}
del3d::CollinearPosition __wrap__ZN5del3d18collinear_positionEPKdS1_S1_(D p, D q, D r) {
    const del3d::CollinearPosition s = __real__ZN5del3d18collinear_positionEPKdS1_S1_(p, q, r);
    record(kPosition, static_cast<int>(s), {p, q, r});
    return s;
}
Sign __wrap__ZN5del3d18coplanar_in_circleEPKdS1_S1_S1_(D a, D b, D c, D p) {
    const Sign s = __real__ZN5del3d18coplanar_in_circleEPKdS1_S1_S1_(a, b, c, p);
    record(kInCircle, s, {a, b, c, p});
    return s;
}
Sign __wrap__ZN5del3d30coplanar_in_circle_unperturbedEPKdS1_S1_S1_(D a, D b, D c, D p) {
    const Sign s = __real__ZN5del3d30coplanar_in_circle_unperturbedEPKdS1_S1_S1_(a, b, c, p);
    record(kInCircleRaw, s, {a, b, c, p});
    return s;
}
}

int main(int argc, char** argv) {
    std::vector<del3d::Point3> pts;
    const char* out_path = nullptr;
    if (argc == 6 && std::strcmp(argv[1], "lattice") == 0) {
        const int nx = std::atoi(argv[2]), ny = std::atoi(argv[3]), nz = std::atoi(argv[4]);
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j)
                for (int k = 0; k < nz; ++k) pts.push_back({double(i), double(j), double(k)});
        out_path = argv[5];
    } else if (argc == 4) {
        const int64_t n = std::atoll(argv[1]);
        const uint32_t seed = static_cast<uint32_t>(std::strtoul(argv[2], nullptr, 10));
        pts.resize(static_cast<size_t>(n));
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
        out_path = argv[3];
    } else {
        std::fprintf(stderr, "usage: %s N SEED OUT | %s lattice NX NY NZ OUT\n", argv[0], argv[0]);
        return 2;
    }
    const int64_t n = static_cast<int64_t>(pts.size());
    g_n = n;
    g_base = &pts[0].x;
    del3d::Delaunay d;
    d.build(pts);

    FILE* f = std::fopen(out_path, "wb");
    if (!f) return 1;
    const uint32_t version = 2;
    const int64_t nrec = static_cast<int64_t>(g_op.size());
    std::fwrite("D3TR", 1, 4, f);
    std::fwrite(&version, 4, 1, f);
    std::fwrite(&n, 8, 1, f);
    std::fwrite(&nrec, 8, 1, f);
    std::fwrite(pts.data(), sizeof(del3d::Point3), pts.size(), f);
    std::fwrite(g_op.data(), 1, g_op.size(), f);
    std::fwrite(g_res.data(), 1, g_res.size(), f);
    std::fwrite(g_idx.data(), 4, g_idx.size(), f);
    std::fclose(f);
    std::printf("points %lld records %lld cells %zu\n", static_cast<long long>(n),
                static_cast<long long>(nrec), d.finite_cells().size());
    return 0;
}
