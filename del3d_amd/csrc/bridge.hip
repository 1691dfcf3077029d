// bridge.hip - conflict regions of an incremental Delaunay insertion with the
// conflict tests batched per breadth-first wave (del3d_bridge_*).
//
// The reference's driver collects a point's conflict region by a depth-first
// flood fill that tests one neighbour cell at a time (src/triangulation.cpp:
// 643-689, the tests at :588-629). A test's answer depends only on the cell's
// vertices and the point, never on the fill's marks, so the region can be
// found in two phases:
//   1. waves: from the start cell, every untested neighbour of the cells found
//      in conflict in the previous wave is tested, all of a wave's tests (over
//      every region of the call) in one mixed predicate batch;
//   2. replay: the reference's depth-first order is walked on the host over
//      the recorded answers, which yields its cells and facets in its order.
// The cells tested are the same in both forms: the region's cells other than
// the start cell, and every cell adjacent to the region.
//
// A wave of fewer than gpu_min tests is answered on the host with the
// kernels' own per-query code (host_scalar.cpp), since one device round trip
// costs 10-20 us; gpu_min = 0 sends every wave to the GPU.
#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "common.h"
#include "host_scalar.h"
#include "predicates.h"

namespace d3 {
namespace {

// What a cell's entry in a region's table records.
constexpr uint8_t kTestIn = 1;    // tested: in conflict
constexpr uint8_t kTestOut = 2;   // tested: not in conflict
constexpr uint8_t kMarkIn = 4;    // replay: reached and in the region
constexpr uint8_t kMarkOut = 8;   // replay: reached and on the boundary

// How a test's answer becomes "in conflict" (src/triangulation.cpp:588-629).
enum Kind : uint8_t {
    kSphere,      // finite cell: in_sphere(v0..v3, p) == +1
    kHullFacet,   // infinite cell, dimension 3: orient3d(a, b, d, p), then on 0
                  // coplanar_in_circle(a, b, d, p) == +1
    kCircle,      // finite triangle, dimension 2: coplanar_in_circle == +1
    kHullEdge,    // infinite triangle, dimension 2: coplanar_orientation(v1, v2, p),
                  // then on 0 collinear_position(v1, p, v2) == middle
};

constexpr int kCollinearMiddle = 2;  // CollinearPosition's "middle" in del3d_predicates2's encoding

// For an infinite cell whose infinite vertex sits in slot i, the slots of the
// hull facet's three vertices in the order the conflict orientation takes
// them (src/triangulation.cpp:592): with p last, the orientation equals that
// of the cell with p in slot i (an even permutation of it), so POSITIVE means
// p lies beyond the facet.
constexpr int kFacetSlots[4][3] = {{2, 1, 3}, {2, 3, 0}, {1, 0, 3}, {0, 1, 2}};

struct Test {
    int32_t region;
    int32_t cell;
    Kind kind;
    int32_t pts[5];  // the first query's points
};

struct Region {
    const del3d_tds_view* view;
    int32_t start;
    int32_t point;                              // global point index of p
    std::unordered_map<int32_t, uint8_t> tab;   // cell -> kTest* | kMark*
    std::vector<int32_t> frontier;
};

}  // namespace
}  // namespace d3

struct del3d_bridge_s {
    std::mutex mu;
    int dev = 0;
    hipStream_t s = nullptr;
    const double* h_xyz = nullptr;
    int64_t n_pts = 0;
    int64_t gpu_min = 0;
    double* d_xyz = nullptr;  // uploaded at the first wave that goes to the GPU
    size_t cap = 0;           // queries the wave buffers hold
    // a wave's records in one block, copied in one transfer: the index
    // records (20 B per query) followed by the op bytes
    uint8_t* h_rec = nullptr;
    int8_t* h_out = nullptr;
    uint8_t* d_rec = nullptr;
    int8_t* d_out = nullptr;
    uint32_t* d_wide = nullptr;  // the one-pass kernel's wide queue
    d3::PredCounters* d_ctr = nullptr;
    int64_t info[8] = {};
};

namespace d3 {
namespace {

void free_wave_buffers(del3d_bridge_s* b) {
    if (b->h_rec) (void)hipHostFree(b->h_rec);
    if (b->h_out) (void)hipHostFree(b->h_out);
    if (b->d_rec) (void)hipFree(b->d_rec);
    if (b->d_out) (void)hipFree(b->d_out);
    if (b->d_wide) (void)hipFree(b->d_wide);
    if (b->d_ctr) (void)hipFree(b->d_ctr);
    b->h_rec = nullptr;
    b->h_out = nullptr;
    b->d_rec = nullptr;
    b->d_out = nullptr;
    b->d_wide = nullptr;
    b->d_ctr = nullptr;
    b->cap = 0;
}

void reserve_wave(del3d_bridge_s* b, size_t n) {
    if (n <= b->cap) return;
    size_t cap = std::max<size_t>(4096, b->cap);
    while (cap < n) cap *= 2;
    D3_HIP(hipStreamSynchronize(b->s));
    free_wave_buffers(b);
    D3_HIP(hipHostMalloc((void**)&b->h_rec, cap * 21, hipHostMallocDefault));
    D3_HIP(hipHostMalloc((void**)&b->h_out, cap, hipHostMallocDefault));
    D3_HIP(hipMalloc((void**)&b->d_rec, cap * 21));
    D3_HIP(hipMalloc((void**)&b->d_out, cap));
    D3_HIP(hipMalloc((void**)&b->d_wide, cap * 4));
    D3_HIP(hipMalloc((void**)&b->d_ctr, sizeof(PredCounters)));
    b->cap = cap;
}

// Answers queries (op as in del3d_predicates_mixed, five point indices each):
// on the host below gpu_min queries, else as one mixed batch on the GPU.
void answer(del3d_bridge_s* b, const std::vector<uint8_t>& op, const std::vector<int32_t>& idx,
            std::vector<int8_t>& out) {
    const size_t n = op.size();
    out.assign(n, 0);
    if (!n) return;
    b->info[2] += (int64_t)n;
    if ((int64_t)n < b->gpu_min) {
        b->info[4] += (int64_t)n;
        for (size_t q = 0; q < n; ++q) {
            P3 p[5];
            int32_t id[5];
            for (int k = 0; k < 5; ++k) {
                const int32_t i = idx[5 * q + k];
                id[k] = i;
                p[k] = P3{b->h_xyz[3 * (int64_t)i], b->h_xyz[3 * (int64_t)i + 1],
                          b->h_xyz[3 * (int64_t)i + 2]};
            }
            out[q] = (int8_t)(op[q] < DEL3D_MIXED_OP2_BASE
                                  ? host_pred3(op[q], p, id)
                                  : host_pred2((uint8_t)(op[q] - DEL3D_MIXED_OP2_BASE), p, id));
        }
        return;
    }
    b->info[1] += 1;
    b->info[3] += (int64_t)n;
    if (!b->d_xyz) {
        D3_HIP(hipMalloc((void**)&b->d_xyz, (size_t)b->n_pts * 24));
        copy_h2d(b->d_xyz, b->h_xyz, (size_t)b->n_pts * 24, b->s);
    }
    reserve_wave(b, n);
    std::memcpy(b->h_rec, idx.data(), n * 20);
    std::memcpy(b->h_rec + n * 20, op.data(), n);
    bool has3 = false, has2 = false;
    for (uint8_t o : op) (o < DEL3D_MIXED_OP2_BASE ? has3 : has2) = true;
    D3_HIP(hipMemcpyAsync(b->d_rec, b->h_rec, n * 21, hipMemcpyHostToDevice, b->s));
    const int32_t* d_idx = reinterpret_cast<const int32_t*>(b->d_rec);
    const uint8_t* d_op = b->d_rec + n * 20;
    if (!has2) {
        // 3D tests only (every wave of a 3D triangulation but the hull's
        // rare follow-ups): the one-pass kernel
        launch_predicates_small(b->d_xyz, b->n_pts, d_op, d_idx, (int64_t)n, b->d_out, b->d_ctr, b->d_wide, b->s);
    } else {
        // the 3D pass writes every query (0 for coplanar op bytes); the
        // coplanar pass then writes only its own ops
        if (has3) launch_predicates(b->d_xyz, b->n_pts, d_op, d_idx, (int64_t)n, b->d_out, b->s, nullptr);
        launch_predicates2(b->d_xyz, b->n_pts, d_op, d_idx, (int64_t)n, b->d_out, b->s, DEL3D_MIXED_OP2_BASE);
    }
    D3_HIP(hipMemcpyAsync(b->h_out, b->d_out, n, hipMemcpyDeviceToHost, b->s));
    D3_HIP(hipStreamSynchronize(b->s));
    std::memcpy(out.data(), b->h_out, n);
}

// Checked accessors of a view; a bad index raises DEL3D_ERR_INVALID.
struct BadView {
    const char* what;
    int64_t value;
};

inline const int32_t* cell_vertices(const del3d_tds_view& v, int32_t c) {
    if (c < 0 || c >= v.n_cells) throw BadView{"cell index outside [0, n_cells)", c};
    return v.cell_v + (int64_t)c * v.cell_stride;
}

inline int32_t neighbour(const del3d_tds_view& v, int32_t c, int i) {
    if (c < 0 || c >= v.n_cells) throw BadView{"cell index outside [0, n_cells)", c};
    return v.cell_n[(int64_t)c * v.cell_stride + i];
}

inline int32_t point_of(const del3d_tds_view& v, int32_t u, int64_t n_pts) {
    if (u < 0 || u >= v.n_vertices) throw BadView{"vertex index outside [0, n_vertices)", u};
    const int64_t p = (int64_t)v.vertex_point[(int64_t)u * v.vertex_stride] + v.point_base;
    if (p < 0 || p >= n_pts) throw BadView{"vertex's point index outside the bridge's points", p};
    return (int32_t)p;
}

// The conflict test of cell c against region r's point (src/triangulation.cpp:
// 588-629), as its first query.
Test make_test(const Region& r, int32_t ri, int32_t c, int64_t n_pts) {
    const del3d_tds_view& v = *r.view;
    const int32_t* cv = cell_vertices(v, c);
    Test t{ri, c, kSphere, {r.point, r.point, r.point, r.point, r.point}};
    if (v.dimension == 3) {
        for (int i = 0; i < 4; ++i) {
            if (cv[i] != v.infinite) continue;
            t.kind = kHullFacet;
            for (int k = 0; k < 3; ++k) t.pts[k] = point_of(v, cv[kFacetSlots[i][k]], n_pts);
            t.pts[3] = r.point;
            return t;
        }
        for (int k = 0; k < 4; ++k) t.pts[k] = point_of(v, cv[k], n_pts);
        t.pts[4] = r.point;
        return t;
    }
    int inf = -1;
    for (int i = 0; i < 4; ++i)
        if (cv[i] == v.infinite) {
            inf = i;
            break;
        }
    if (inf < 0) {
        t.kind = kCircle;
        for (int k = 0; k < 3; ++k) t.pts[k] = point_of(v, cv[k], n_pts);
        t.pts[3] = r.point;
        return t;
    }
    if (inf > 2) throw BadView{"infinite vertex in slot 3 of a dimension-2 cell", c};
    // the finite vertices counterclockwise after the infinite one
    // (src/tds.h:59-65: ccw(i) = (i+1) mod 3, cw(i) = (i+2) mod 3)
    t.kind = kHullEdge;
    t.pts[0] = point_of(v, cv[(inf + 1) % 3], n_pts);
    t.pts[1] = point_of(v, cv[(inf + 2) % 3], n_pts);
    t.pts[2] = r.point;
    return t;
}

void first_query(const Test& t, uint8_t* op, int32_t* idx) {
    switch (t.kind) {
    case kSphere: *op = DEL3D_OP_IN_SPHERE; break;
    case kHullFacet: *op = DEL3D_OP_ORIENT3D; break;
    case kCircle: *op = DEL3D_MIXED_OP2_BASE + DEL3D_OP2_IN_CIRCLE; break;
    case kHullEdge: *op = DEL3D_MIXED_OP2_BASE + DEL3D_OP2_COPLANAR_ORIENTATION; break;
    }
    std::memcpy(idx, t.pts, 20);
}

// The follow-up of a hull test whose orientation was 0.
void second_query(const Test& t, uint8_t* op, int32_t* idx) {
    if (t.kind == kHullFacet) {
        *op = DEL3D_MIXED_OP2_BASE + DEL3D_OP2_IN_CIRCLE;  // coplanar_in_circle(a, b, d, p)
        std::memcpy(idx, t.pts, 20);
    } else {
        *op = DEL3D_MIXED_OP2_BASE + DEL3D_OP2_COLLINEAR_POSITION;  // collinear_position(v1, p, v2)
        const int32_t v1 = t.pts[0], v2 = t.pts[1], p = t.pts[2];
        const int32_t q[5] = {v1, p, v2, v2, v2};
        std::memcpy(idx, q, 20);
    }
}

int run_regions(del3d_bridge_s* b, int32_t n_regions, const del3d_tds_view* views,
                const int32_t* starts, const int32_t* points, int32_t* cells_out, int64_t cells_cap,
                int64_t* cells_off, int32_t* facets_out, int64_t facets_cap, int64_t* facets_off) {
    std::vector<Region> rs((size_t)n_regions);
    for (int32_t ri = 0; ri < n_regions; ++ri) {
        const del3d_tds_view& v = views[ri];
        if (v.dimension != 2 && v.dimension != 3)
            return set_error(DEL3D_ERR_INVALID, "region %d: dimension %d (2 or 3 only)", ri, v.dimension);
        if (!v.cell_v || !v.cell_n || !v.vertex_point || v.cell_stride < 4 || v.vertex_stride < 1 ||
            v.n_cells <= 0 || v.n_vertices <= 0)
            return set_error(DEL3D_ERR_INVALID, "region %d: incomplete view", ri);
        if (points[ri] < 0 || points[ri] >= b->n_pts)
            return set_error(DEL3D_ERR_INVALID, "region %d: point %d outside the bridge's points", ri,
                             points[ri]);
        if (starts[ri] < 0 || starts[ri] >= v.n_cells)
            return set_error(DEL3D_ERR_INVALID, "region %d: start cell %d outside the view", ri, starts[ri]);
        Region& r = rs[(size_t)ri];
        r.view = &v;
        r.start = starts[ri];
        r.point = points[ri];
        r.tab.reserve(128);
        r.tab[r.start] = kTestIn;  // the located cell is in conflict (it is not tested)
        r.frontier.push_back(r.start);
    }
    // phase 1: breadth-first waves, each answered by one batch over all regions
    std::vector<Test> tests;
    std::vector<uint8_t> op, op2;
    std::vector<int32_t> idx, idx2;
    std::vector<int8_t> res, res2;
    std::vector<size_t> follow;
    for (;;) {
        tests.clear();
        for (int32_t ri = 0; ri < n_regions; ++ri) {
            Region& r = rs[(size_t)ri];
            const int nnb = r.view->dimension + 1;
            for (int32_t c : r.frontier)
                for (int i = 0; i < nnb; ++i) {
                    const int32_t t = neighbour(*r.view, c, i);
                    auto ins = r.tab.emplace(t, 0);
                    if (!ins.second) continue;  // tested already, or queued in this wave
                    ins.first->second = kTestOut;  // placeholder until answered
                    tests.push_back(make_test(r, ri, t, b->n_pts));
                }
            r.frontier.clear();
        }
        if (tests.empty()) break;
        b->info[0] += 1;
        op.resize(tests.size());
        idx.resize(tests.size() * 5);
        for (size_t k = 0; k < tests.size(); ++k) first_query(tests[k], &op[k], &idx[5 * k]);
        answer(b, op, idx, res);
        follow.clear();
        for (size_t k = 0; k < tests.size(); ++k)
            if ((tests[k].kind == kHullFacet || tests[k].kind == kHullEdge) && res[k] == 0) follow.push_back(k);
        if (!follow.empty()) {
            b->info[5] += (int64_t)follow.size();
            op2.resize(follow.size());
            idx2.resize(follow.size() * 5);
            for (size_t j = 0; j < follow.size(); ++j) second_query(tests[follow[j]], &op2[j], &idx2[5 * j]);
            answer(b, op2, idx2, res2);
        }
        size_t j = 0;
        for (size_t k = 0; k < tests.size(); ++k) {
            const Test& t = tests[k];
            bool in;
            if (j < follow.size() && follow[j] == k) {
                in = t.kind == kHullFacet ? res2[j] == 1 : res2[j] == kCollinearMiddle;
                ++j;
            } else {
                in = res[k] == 1;
            }
            Region& r = rs[(size_t)t.region];
            r.tab[t.cell] = in ? kTestIn : kTestOut;
            if (in) r.frontier.push_back(t.cell);
        }
    }
    // phase 2: the reference's depth-first order over the recorded answers
    // (src/triangulation.cpp:668-688)
    int64_t nc = 0, nf = 0;
    bool over = false;
    std::vector<int32_t> stack;
    for (int32_t ri = 0; ri < n_regions; ++ri) {
        Region& r = rs[(size_t)ri];
        const int nnb = r.view->dimension + 1;
        cells_off[ri] = nc;
        facets_off[ri] = nf;
        auto push_cell = [&](int32_t c) {
            if (nc < cells_cap) cells_out[nc] = c;
            else over = true;
            ++nc;
        };
        stack.clear();
        stack.push_back(r.start);
        r.tab[r.start] |= kMarkIn;
        push_cell(r.start);
        while (!stack.empty()) {
            const int32_t c = stack.back();
            stack.pop_back();
            for (int i = 0; i < nnb; ++i) {
                const int32_t t = neighbour(*r.view, c, i);
                uint8_t& e = r.tab.at(t);
                if (e & kMarkIn) continue;
                if (!(e & kMarkOut)) {
                    if (e & kTestIn) {
                        stack.push_back(t);
                        e |= kMarkIn;
                        push_cell(t);
                        continue;
                    }
                    e |= kMarkOut;
                }
                if (nf < facets_cap) {
                    facets_out[2 * nf] = c;
                    facets_out[2 * nf + 1] = i;
                } else {
                    over = true;
                }
                ++nf;
            }
        }
    }
    cells_off[n_regions] = nc;
    facets_off[n_regions] = nf;
    if (over)
        return set_error(DEL3D_ERR_INVALID,
                         "output capacity: %lld cells and %lld facets needed (cells_off[n] and "
                         "facets_off[n] hold the totals)",
                         (long long)nc, (long long)nf);
    return DEL3D_OK;
}

}  // namespace
}  // namespace d3

using namespace d3;

extern "C" {

int del3d_bridge_create(const double* xyz, int64_t n_pts, int64_t gpu_min_queries, del3d_bridge* bridge_out) {
    return guarded([&] {
        if (!bridge_out) return set_error(DEL3D_ERR_INVALID, "bridge_out is null");
        *bridge_out = nullptr;
        if (n_pts <= 0 || !xyz) return set_error(DEL3D_ERR_INVALID, "no points");
        if (n_pts > 0x7FFFFFFF) return set_error(DEL3D_ERR_INVALID, "more than 2^31-1 points");
        if (gpu_min_queries < 0) return set_error(DEL3D_ERR_INVALID, "negative gpu_min_queries");
        auto* b = new del3d_bridge_s;
        b->h_xyz = xyz;
        b->n_pts = n_pts;
        b->gpu_min = gpu_min_queries;
        if (hipGetDevice(&b->dev) != hipSuccess || hipStreamCreateWithFlags(&b->s, hipStreamNonBlocking) != hipSuccess) {
            (void)hipGetLastError();
            b->s = nullptr;
            if (gpu_min_queries != INT64_MAX) {
                delete b;
                return set_error(DEL3D_ERR_DEVICE, "no usable device for the bridge's GPU waves");
            }
        }
        *bridge_out = b;
        return DEL3D_OK;
    });
}

int del3d_bridge_conflicts(del3d_bridge b, int32_t n_regions, const del3d_tds_view* views,
                           const int32_t* start_cells, const int32_t* points, int32_t* cells_out,
                           int64_t cells_cap, int64_t* cells_off, int32_t* facets_out,
                           int64_t facets_cap, int64_t* facets_off) {
    return guarded([&] {
        if (!b) return set_error(DEL3D_ERR_INVALID, "null bridge");
        if (n_regions < 0) return set_error(DEL3D_ERR_INVALID, "negative n_regions");
        if (n_regions > 0 && (!views || !start_cells || !points || !cells_off || !facets_off))
            return set_error(DEL3D_ERR_INVALID, "null argument");
        if (cells_cap < 0 || facets_cap < 0 || (cells_cap > 0 && !cells_out) || (facets_cap > 0 && !facets_out))
            return set_error(DEL3D_ERR_INVALID, "bad output buffer");
        if (n_regions == 0) {
            if (cells_off) cells_off[0] = 0;
            if (facets_off) facets_off[0] = 0;
            return DEL3D_OK;
        }
        std::lock_guard<std::mutex> lock(b->mu);
        // the waves run on the bridge's device; the caller's current device is
        // restored afterwards
        struct DeviceScope {
            int prev = -1;
            explicit DeviceScope(int dev) {
                if (hipGetDevice(&prev) != hipSuccess) prev = -1;
                if (prev != dev) D3_HIP(hipSetDevice(dev));
            }
            ~DeviceScope() {
                int cur = -1;
                if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
            }
        };
        std::unique_ptr<DeviceScope> scope;
        if (b->s) scope.reset(new DeviceScope(b->dev));
        try {
            return run_regions(b, n_regions, views, start_cells, points, cells_out, cells_cap, cells_off,
                               facets_out, facets_cap, facets_off);
        } catch (const BadView& e) {
            return set_error(DEL3D_ERR_INVALID, "bad view: %s (%lld)", e.what, (long long)e.value);
        }
    });
}

int del3d_bridge_info(del3d_bridge b, int64_t* info) {
    return guarded([&] {
        if (!b || !info) return set_error(DEL3D_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(b->mu);
        for (int k = 0; k < 6; ++k) info[k] = b->info[k];
        return DEL3D_OK;
    });
}

int del3d_bridge_destroy(del3d_bridge b) {
    return guarded([&] {
        if (!b) return DEL3D_OK;
        if (b->s) {
            (void)hipSetDevice(b->dev);
            (void)hipStreamSynchronize(b->s);
            free_wave_buffers(b);
            if (b->d_xyz) (void)hipFree(b->d_xyz);
            (void)hipStreamDestroy(b->s);
        }
        delete b;
        return DEL3D_OK;
    });
}

}  // extern "C"
