// ref_shim.cpp - plain C entry points over the compiled reference.
//
// TEST INFRASTRUCTURE ONLY. Compiled by oracle/Makefile together with the
// reference's own, unchanged sources (read where they lie in the del3d source
// tree; nothing of them is copied here) into oracle/_ref/libdel3d_ref.so, so
// that tests can compare against the real reference and bench.py can time it
// as the CPU baseline. The declarations used are the reference's public ones:
//   include/del3d/predicates.h:26,49-50,54-55  orient3d / in_sphere(_unperturbed)
//   include/del3d/predicates2.h:26-69          the coplanar predicates
//   include/del3d/spatial_sort.h:31,37         Point3 / spatial_sort
//   include/del3d/delaunay.h:29-49             circumcenter / circumcenter_lazy
//   include/del3d/delaunay.h:62-92             Delaunay (build, finite_cells)
#include "del3d/delaunay.h"
#include "del3d/predicates.h"
#include "del3d/predicates2.h"
#include "del3d/spatial_sort.h"

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

extern "C" {

int ref_orient3d(const double* a, const double* b, const double* c, const double* d) {
    return static_cast<int>(del3d::orient3d(a, b, c, d));
}

int ref_in_sphere(const double* a, const double* b, const double* c, const double* d,
                  const double* e) {
    return static_cast<int>(del3d::in_sphere(a, b, c, d, e));
}

int ref_in_sphere_unperturbed(const double* a, const double* b, const double* c,
                              const double* d, const double* e) {
    return static_cast<int>(del3d::in_sphere_unperturbed(a, b, c, d, e));
}

// op 0 = orient3d(r0..r3), 1 = in_sphere(r0..r4), 2 = in_sphere_unperturbed.
void ref_pred_batch(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx5,
                    int64_t n_q, int8_t* out) {
    (void)n_pts;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const double* p0 = xyz + 3 * int64_t(r[0]);
        const double* p1 = xyz + 3 * int64_t(r[1]);
        const double* p2 = xyz + 3 * int64_t(r[2]);
        const double* p3 = xyz + 3 * int64_t(r[3]);
        int s;
        if (op[q] == 0) {
            s = del3d::orient3d(p0, p1, p2, p3);
        } else {
            const double* p4 = xyz + 3 * int64_t(r[4]);
            s = op[q] == 1 ? del3d::in_sphere(p0, p1, p2, p3, p4)
                           : del3d::in_sphere_unperturbed(p0, p1, p2, p3, p4);
        }
        out[q] = static_cast<int8_t>(s);
    }
}

// Coplanar batch: op 0 coplanar_orientation(p,q,r), 1 coplanar_orientation(p,q,r,s),
// 2 collinear, 3 equal(p,q), 4 collinear_position, 5 coplanar_in_circle,
// 6 coplanar_in_circle_unperturbed, 7 orient2d on the xy coordinates of (p,q,r).
// Points are pointers into xyz, so equal indices are the same pointer. Op 8
// (the driver's private plain-double orientation) is not reachable from here:
// it gets 0 (libdel3d_ref_driver.so answers it).
void ref_pred2_batch(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx5,
                     int64_t n_q, int8_t* out) {
    (void)n_pts;
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        const int np = op[q] == 3 ? 2 : (op[q] == 1 || op[q] == 5 || op[q] == 6 || op[q] == 8) ? 4 : 3;
        const double* p[4];
        for (int k = 0; k < 4; ++k) p[k] = k < np ? xyz + 3 * int64_t(r[k]) : xyz;
        int s = 0;
        switch (op[q]) {
            case 0: s = del3d::coplanar_orientation(p[0], p[1], p[2]); break;
            case 1: s = del3d::coplanar_orientation(p[0], p[1], p[2], p[3]); break;
            case 2: s = del3d::collinear(p[0], p[1], p[2]) ? 1 : 0; break;
            case 3: s = del3d::equal(p[0], p[1]) ? 1 : 0; break;
            case 4: s = static_cast<int>(del3d::collinear_position(p[0], p[1], p[2])); break;
            case 5: s = del3d::coplanar_in_circle(p[0], p[1], p[2], p[3]); break;
            case 6: s = del3d::coplanar_in_circle_unperturbed(p[0], p[1], p[2], p[3]); break;
            case 7: s = del3d::orient2d(p[0][0], p[0][1], p[1][0], p[1][1], p[2][0], p[2][1]); break;
            default: break;
        }
        out[q] = static_cast<int8_t>(s);
    }
}

// Circumcentres of cells (4 point indices each): lazy != 0 selects
// circumcenter_lazy (include/del3d/delaunay.h:48-49), else circumcenter (:29-30).
void ref_circumcenters(const double* xyz, const int32_t* cells, int64_t n, int lazy, double* out) {
    for (int64_t c = 0; c < n; ++c) {
        del3d::Point3 p[4];
        for (int k = 0; k < 4; ++k) {
            const double* v = xyz + 3 * int64_t(cells[4 * c + k]);
            p[k] = {v[0], v[1], v[2]};
        }
        const del3d::Point3 r = lazy ? del3d::circumcenter_lazy(p[0], p[1], p[2], p[3])
                                     : del3d::circumcenter(p[0], p[1], p[2], p[3]);
        out[3 * c] = r.x;
        out[3 * c + 1] = r.y;
        out[3 * c + 2] = r.z;
    }
}

// The same batch split over `threads` std::threads (independent queries).
void ref_pred_batch_mt(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx5,
                       int64_t n_q, int8_t* out, int threads) {
    if (threads <= 1) { ref_pred_batch(xyz, n_pts, op, idx5, n_q, out); return; }
    std::vector<std::thread> pool;
    const int64_t chunk = (n_q + threads - 1) / threads;
    for (int t = 0; t < threads; ++t) {
        const int64_t b = t * chunk, e = b + chunk < n_q ? b + chunk : n_q;
        if (b >= e) break;
        pool.emplace_back([=] { ref_pred_batch(xyz, n_pts, op + b, idx5 + 5 * b, e - b, out + b); });
    }
    for (auto& th : pool) th.join();
}

void ref_spatial_sort(const double* xyz, int64_t n, int32_t* order) {
    std::vector<del3d::Point3> pts(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) pts[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    std::vector<int> ord(order, order + n);
    del3d::spatial_sort(pts, ord);
    std::memcpy(order, ord.data(), sizeof(int32_t) * static_cast<size_t>(n));
}

// Builds the triangulation and returns the number of finite cells; writes up
// to `cap` of them (4 ints each) into cells_out, in the library's cell order.
int64_t ref_delaunay_cells(const double* xyz, int64_t n, int32_t* cells_out, int64_t cap) {
    std::vector<del3d::Point3> pts(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) pts[i] = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
    del3d::Delaunay d;
    d.build(pts);
    const auto& cells = d.finite_cells();
    const int64_t m = static_cast<int64_t>(cells.size());
    for (int64_t i = 0; i < m && i < cap; ++i)
        for (int k = 0; k < 4; ++k) cells_out[4 * i + k] = cells[i][k];
    return m;
}

}  // extern "C"
