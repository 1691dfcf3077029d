// bridge_delaunay.cpp - the reference's unchanged Delaunay driver with its
// conflict-region flood fill answered by the library's wave bridge
// (del3d_bridge_conflicts), checked against the same driver's own build.
//
// TEST INFRASTRUCTURE. The driver's Builder is a class internal to
// src/triangulation.cpp, so this unit includes the unchanged source file (the
// reference tree is on the include path, as for ref_driver.cpp) with its
// members made accessible. oracle/Makefile links it against libdel3d_amd.so,
// which serves the driver's other predicate calls and its spatial sort (the
// drop-in), and builds oracle/_ref/dropin_bridge_delaunay.
//
// The bridged insertion restates Builder::insert (src/triangulation.cpp:
// 702-770) for dimensions 2 and 3, with the call of find_conflicts (:643-689)
// replaced by del3d_bridge_conflicts plus the marks the star rebuild expects
// (the region's cells InConflict, the cells across its facets OnBoundary);
// below dimension 2 the reference's own insert runs. The build loop restates
// Builder::build (:62-79) and the output walk Delaunay::build (:932-967).
//
//   dropin_bridge_delaunay cases <gpu_min>
//       the reference test's inputs (tests/test_delaunay.cpp:185-243)
//   dropin_bridge_delaunay rand48 <n> <gpu_min> [reps]
//       n points from the rand48 generator seeded 1, coordinates x, y, z =
//       successive states * 2^-48
//   dropin_bridge_delaunay lockstep <k> <n> <gpu_min>
//       k builds of n rand48 points (seed s = 1..k: initial state
//       (s << 16) | 0x330E) advanced together, one bridge call per insertion
//       step for all builds that need a conflict region
// Every case prints one JSON line: whether the bridged build's cells (all
// cells, infinite ones included, in container order) equal the driver's own
// build, the finite cells' hash in the reference test's format, and the
// bridge's counters.
#include <algorithm>
#include <array>
#include <cassert>
#include <cfenv>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#define private public
#include "src/triangulation.cpp"
#undef private

#include "pointsets.h"

#include "del3d_amd.h"

namespace {

using del3d::Builder;
using del3d::Delaunay;
using del3d::Point3;
using del3d::detail::Cell;
using del3d::detail::Index;
using del3d::detail::kNull;
using del3d::detail::Mark;
using del3d::detail::Tds;
using del3d::detail::Vertex;

static_assert(sizeof(Cell) % sizeof(int32_t) == 0, "Cell is read as a strided int32 table");
static_assert(sizeof(Vertex) % sizeof(int32_t) == 0, "Vertex is read as a strided int32 table");
static_assert(sizeof(Point3) == 24, "Point3 is three packed doubles");

void die(const char* what) {
    std::fprintf(stderr, "bridge_delaunay: %s: %s\n", what, del3d_last_error());
    std::exit(2);
}

del3d_tds_view view_of(const Tds& t, int64_t point_base) {
    del3d_tds_view v{};
    v.cell_v = t.cells.items_[0].v.data();
    v.cell_n = t.cells.items_[0].n.data();
    v.cell_stride = sizeof(Cell) / sizeof(int32_t);
    v.n_cells = (int64_t)t.cells.items_.size();
    v.vertex_point = &t.vertices.items_[0].info;
    v.vertex_stride = sizeof(Vertex) / sizeof(int32_t);
    v.n_vertices = (int64_t)t.vertices.items_.size();
    v.point_base = point_base;
    v.infinite = t.infinite;
    v.dimension = t.dimension;
    return v;
}

// One build driven step by step.
struct Build {
    const std::vector<Point3>* pts;
    int64_t base;  // first point of this build in the bridge's points
    Builder b;
    std::vector<int> order;
    Index hint = kNull;
    // the pending insertion's located cell and point
    bool need = false;
    Index cell = kNull;
    int idx = -1;
    std::vector<Index> conflict;
    std::vector<std::pair<Index, int>> facets;

    Build(const std::vector<Point3>& p, int64_t base_) : pts(&p), base(base_), b(p) {
        order.resize(p.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = int(i);
        del3d::spatial_sort(p, order);
    }

    // Builder::insert up to the flood fill. Returns true when the point needs
    // a conflict region; otherwise the insertion is complete.
    bool begin(int i) {
        idx = i;
        need = false;
        Tds& t = b.tds_;
        if (t.dimension < 2) {  // no conflict region below dimension 2
            finish(b.insert(idx, hint));
            return false;
        }
        const double* p = b.pt_of(idx);
        del3d::LocateType lt = del3d::LT_OUTSIDE_AFFINE_HULL;
        int li = -1, lj = -1;
        const Index start = hint == kNull ? b.infinite_cell() : t.vertices[hint].cell;
        const Index c = b.locate(p, lt, li, lj, start);
        if (t.dimension == 2 && lt == del3d::LT_OUTSIDE_AFFINE_HULL) {
            finish(b.insert_outside_affine_hull(p));
            return false;
        }
        if (lt == del3d::LT_VERTEX) {
            finish(t.cells[c].v[li]);
            return false;
        }
        need = true;
        cell = c;
        return true;
    }

    // The star rebuild from the bridge's region (cells and facets in the
    // reference fill's order).
    void end(const int32_t* cells, int64_t nc, const int32_t* fac, int64_t nf) {
        Tds& t = b.tds_;
        conflict.assign(cells, cells + nc);
        facets.clear();
        for (int64_t k = 0; k < nf; ++k) facets.emplace_back(fac[2 * k], fac[2 * k + 1]);
        for (Index c : conflict) t.cells[c].mark = Mark::InConflict;
        for (const auto& f : facets) t.cells[t.cells[f.first].n[f.second]].mark = Mark::OnBoundary;
        const Index v = t.create_vertex();
        if (t.dimension == 3 && Tds::is_small_hole(facets.size()))
            t.insert_in_small_hole(conflict, facets, v);
        else
            t.insert_in_hole(conflict, facets.back().first, facets.back().second, v);
        finish(v);
    }

    void finish(Index v) {
        hint = v;
        if (hint != kNull) b.tds_.vertices[hint].info = idx;
        need = false;
    }

    std::vector<Delaunay::Tet> all_cells() {
        std::vector<Delaunay::Tet> out;
        Tds& t = b.tds_;
        if (t.dimension < 3) return out;
        t.cells.for_each([&](Index c) {
            Delaunay::Tet a;
            for (int i = 0; i < 4; ++i) {
                const Index vv = t.cells[c].v[i];
                a[i] = vv == t.infinite ? -1 : t.vertices[vv].info;
            }
            out.push_back(a);
        });
        return out;
    }
};

// Advances builds together: at step i every build begins its i-th point, one
// bridge call answers every pending region, and every pending build ends it.
void run_lockstep(std::vector<Build*>& builds, del3d_bridge br) {
    size_t steps = 0;
    for (Build* b : builds) steps = std::max(steps, b->order.size());
    std::vector<del3d_tds_view> views;
    std::vector<int32_t> starts, points, cells, facets;
    std::vector<int64_t> coff, foff;
    std::vector<Build*> pend;
    cells.resize(4096);
    facets.resize(2 * 4096);
    for (size_t i = 0; i < steps; ++i) {
        pend.clear();
        views.clear();
        starts.clear();
        points.clear();
        for (Build* b : builds)
            if (i < b->order.size() && b->begin(b->order[i])) {
                pend.push_back(b);
                views.push_back(view_of(b->b.tds_, b->base));
                starts.push_back(b->cell);
                points.push_back((int32_t)(b->base + b->idx));
            }
        if (pend.empty()) continue;
        const int32_t n = (int32_t)pend.size();
        coff.assign((size_t)n + 1, 0);
        foff.assign((size_t)n + 1, 0);
        for (;;) {
            const int rc = del3d_bridge_conflicts(br, n, views.data(), starts.data(), points.data(), cells.data(),
                                                  (int64_t)cells.size(), coff.data(), facets.data(),
                                                  (int64_t)facets.size() / 2, foff.data());
            if (rc == DEL3D_OK) break;
            if (coff[n] <= (int64_t)cells.size() && 2 * foff[n] <= (int64_t)facets.size()) die("bridge");
            cells.resize((size_t)std::max<int64_t>(coff[n], (int64_t)cells.size()));
            facets.resize((size_t)std::max<int64_t>(2 * foff[n], (int64_t)facets.size()));
        }
        for (int32_t r = 0; r < n; ++r)
            pend[r]->end(cells.data() + coff[r], coff[r + 1] - coff[r], facets.data() + 2 * foff[r],
                         foff[r + 1] - foff[r]);
    }
}

uint64_t hash_finite(const std::vector<Delaunay::Tet>& all) {
    // the reference test's format (tests/test_delaunay.cpp:44-56): FNV-1a 64
    // over the cell count, then every vertex index as an int64, 8 bytes each,
    // over the finite cells
    std::vector<Delaunay::Tet> fin;
    for (const auto& t : all)
        if (t[0] >= 0 && t[1] >= 0 && t[2] >= 0 && t[3] >= 0) fin.push_back(t);
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xffull;
            h *= 1099511628211ull;
        }
    };
    mix(fin.size());
    for (const auto& t : fin)
        for (int v : t) mix((uint64_t)(int64_t)v);
    return h;
}

uint64_t hash_timing_format(const std::vector<Delaunay::Tet>& all) {
    // delaunay_timing.cpp's format: FNV-1a 64 over the 4 bytes of every vertex
    // index of the finite cells
    uint64_t h = 1469598103934665603ull;
    for (const auto& t : all) {
        if (t[0] < 0 || t[1] < 0 || t[2] < 0 || t[3] < 0) continue;
        for (int v : t)
            for (int b = 0; b < 4; ++b) {
                h ^= ((uint64_t)(uint32_t)v >> (8 * b)) & 0xffull;
                h *= 1099511628211ull;
            }
    }
    return h;
}

size_t finite_count(const std::vector<Delaunay::Tet>& all) {
    size_t n = 0;
    for (const auto& t : all) n += (t[0] >= 0 && t[1] >= 0 && t[2] >= 0 && t[3] >= 0);
    return n;
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

std::vector<Point3> rand48_points(long n, uint64_t seed) {
    std::vector<Point3> pts((size_t)n);
    uint64_t x = (seed << 16) | 0x330Eull;
    auto next = [&x] {
        x = (0x5DEECE66Dull * x + 0xBull) & ((1ull << 48) - 1);
        return (double)x * 0x1p-48;
    };
    for (auto& p : pts) {
        p.x = next();
        p.y = next();
        p.z = next();
    }
    return pts;
}

// One case: the bridged build and the driver's own build of the same points.
bool run_case(const char* name, const std::vector<Point3>& pts, int64_t gpu_min, int reps) {
    std::vector<Delaunay::Tet> bridged, driver;
    double t_bridged = 1e30, t_driver = 1e30;
    int64_t info[6] = {};
    for (int r = 0; r < reps; ++r) {
        if (pts.empty()) break;
        del3d_bridge br = nullptr;
        if (del3d_bridge_create(&pts[0].x, (int64_t)pts.size(), gpu_min, &br) != DEL3D_OK) die("bridge_create");
        const double t0 = now_s();
        Build b(pts, 0);
        std::vector<Build*> one{&b};
        run_lockstep(one, br);
        bridged = b.all_cells();
        t_bridged = std::min(t_bridged, now_s() - t0);
        if (del3d_bridge_info(br, info) != DEL3D_OK) die("bridge_info");
        del3d_bridge_destroy(br);
        const double t1 = now_s();
        Delaunay d;
        d.build(pts);
        driver = d.all_cells();
        t_driver = std::min(t_driver, now_s() - t1);
    }
    const bool same = bridged == driver;
    std::printf("{\"case\": \"%s\", \"points\": %zu, \"cells\": %zu, \"finite\": %zu, \"hash\": \"0x%016llx\", "
                "\"timing_hash\": \"%016llx\", \"same_as_driver\": %s, \"bridged_s\": %.6f, \"driver_s\": %.6f, \"gpu_min\": %lld, "
                "\"bridge\": {\"waves\": %lld, \"gpu_batches\": %lld, \"tests\": %lld, \"gpu_tests\": %lld, "
                "\"host_tests\": %lld, \"follow_ups\": %lld}}\n",
                name, pts.size(), bridged.size(), finite_count(bridged),
                (unsigned long long)hash_finite(bridged), (unsigned long long)hash_timing_format(bridged),
                same ? "true" : "false", pts.empty() ? 0.0 : t_bridged,
                pts.empty() ? 0.0 : t_driver, (long long)gpu_min, (long long)info[0], (long long)info[1],
                (long long)info[2], (long long)info[3], (long long)info[4], (long long)info[5]);
    std::fflush(stdout);
    return same;
}

std::vector<Point3> to_points(const std::vector<pointsets::Pt>& v) {
    std::vector<Point3> out;
    for (const auto& p : v) out.push_back({p[0], p[1], p[2]});
    return out;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc < 3) {
        std::fprintf(stderr, "usage: %s cases <gpu_min> | rand48 <n> <gpu_min> [reps] | lockstep <k> <n> <gpu_min>\n",
                     argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    bool ok = true;
    if (mode == "cases") {
        const int64_t gpu_min = std::atoll(argv[2]);
        // tests/test_delaunay.cpp:185-243: the degenerate small inputs, random
        // cubes, lattice, plane + apex, box shell and quantised surfaces
        std::vector<Point3> dup(20, Point3{1.0, 2.0, 3.0});
        ok &= run_case("20 identical points", dup, gpu_min, 1);
        std::vector<Point3> line;
        for (int i = 0; i < 10; ++i) line.push_back({double(i), 2.0 * i, 3.0 * i});
        ok &= run_case("collinear", line, gpu_min, 1);
        std::vector<Point3> flat;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) flat.push_back({double(i), double(j), 0.0});
        ok &= run_case("coplanar", flat, gpu_min, 1);
        for (int n : {5, 8, 20, 100, 500}) {
            const std::string name = "random_cube_seed4 " + std::to_string(n);
            ok &= run_case(name.c_str(), to_points(pointsets::random_cube((size_t)n, 4u)), gpu_min, 1);
        }
        ok &= run_case("lattice_5_5_5", to_points(pointsets::lattice(5, 5, 5)), gpu_min, 1);
        std::vector<Point3> plane;
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) plane.push_back({double(i), double(j), 0.0});
        plane.push_back({2.0, 2.0, 3.0});
        ok &= run_case("plane6x6_plus_apex_2_2_3", plane, gpu_min, 1);
        ok &= run_case("box_shell_5", to_points(pointsets::box_shell(5)), gpu_min, 1);
        ok &= run_case("cylinder_300_q32", to_points(pointsets::cylinder(300)), gpu_min, 1);
        ok &= run_case("sphere_300_q32", to_points(pointsets::sphere(300)), gpu_min, 1);
        ok &= run_case("torus_300_q32", to_points(pointsets::torus(300)), gpu_min, 1);
        ok &= run_case("lattice_12_12_12", to_points(pointsets::lattice(12, 12, 12)), gpu_min, 1);
    } else if (mode == "rand48" && argc >= 4) {
        const long n = std::atol(argv[2]);
        const int64_t gpu_min = std::atoll(argv[3]);
        const int reps = argc > 4 ? std::atoi(argv[4]) : 1;
        const std::string name = "rand48 " + std::to_string(n);
        ok &= run_case(name.c_str(), rand48_points(n, 1), gpu_min, reps);
    } else if (mode == "lockstep" && argc >= 5) {
        const int k = std::atoi(argv[2]);
        const long n = std::atol(argv[3]);
        const int64_t gpu_min = std::atoll(argv[4]);
        std::vector<std::vector<Point3>> sets;
        std::vector<Point3> all;
        for (int s = 0; s < k; ++s) {
            sets.push_back(rand48_points(n, (uint64_t)s + 1));
            all.insert(all.end(), sets.back().begin(), sets.back().end());
        }
        del3d_bridge br = nullptr;
        if (del3d_bridge_create(&all[0].x, (int64_t)all.size(), gpu_min, &br) != DEL3D_OK) die("bridge_create");
        const double t0 = now_s();
        std::vector<Build> builds;
        builds.reserve((size_t)k);
        for (int s = 0; s < k; ++s) builds.emplace_back(sets[(size_t)s], (int64_t)s * n);
        std::vector<Build*> ptrs;
        for (auto& b : builds) ptrs.push_back(&b);
        run_lockstep(ptrs, br);
        const double t_bridged = now_s() - t0;
        int64_t info[6] = {};
        if (del3d_bridge_info(br, info) != DEL3D_OK) die("bridge_info");
        del3d_bridge_destroy(br);
        int same = 0;
        const double t1 = now_s();
        std::vector<std::vector<Delaunay::Tet>> drv((size_t)k);
        for (int s = 0; s < k; ++s) {
            Delaunay d;
            d.build(sets[(size_t)s]);
            drv[(size_t)s] = d.all_cells();
        }
        const double t_driver = now_s() - t1;
        for (int s = 0; s < k; ++s) same += builds[(size_t)s].all_cells() == drv[(size_t)s];
        ok = same == k;
        std::printf("{\"case\": \"lockstep %d x %ld\", \"builds\": %d, \"same_as_driver\": %d, \"bridged_s\": %.6f, "
                    "\"driver_s\": %.6f, \"gpu_min\": %lld, \"bridge\": {\"waves\": %lld, \"gpu_batches\": %lld, "
                    "\"tests\": %lld, \"gpu_tests\": %lld, \"host_tests\": %lld, \"follow_ups\": %lld}}\n",
                    k, n, k, same, t_bridged, t_driver, (long long)gpu_min, (long long)info[0], (long long)info[1],
                    (long long)info[2], (long long)info[3], (long long)info[4], (long long)info[5]);
    } else {
        std::fprintf(stderr, "bad arguments\n");
        return 2;
    }
    return ok ? 0 : 1;
}
