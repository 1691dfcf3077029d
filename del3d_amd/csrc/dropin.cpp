// dropin.cpp - the reference's C++ symbols over the C ABI.
//
// Scalar predicate calls, 3D and coplanar, run one query on the GPU
// (del3d_orient3d & co.);
// spatial_sort runs the GPU sort on a host copy of the data. The reference's
// signatures have no error channel, so a device failure aborts with the
// library's message rather than returning a sign that was never computed.
#include <cstdio>
#include <cstdlib>

#include "../../include/del3d_amd.h"
#include "../../include/del3d_amd_dropin.hpp"

namespace del3d {

Sign orient3d(const double* a, const double* b, const double* c, const double* d) {
    return static_cast<Sign>(del3d_orient3d(a, b, c, d));
}

Sign in_sphere(const double* a, const double* b, const double* c, const double* d, const double* e) {
    return static_cast<Sign>(del3d_in_sphere(a, b, c, d, e));
}

Sign in_sphere_unperturbed(const double* a, const double* b, const double* c, const double* d,
                           const double* e) {
    return static_cast<Sign>(del3d_in_sphere_unperturbed(a, b, c, d, e));
}

Sign orient2d(double px, double py, double qx, double qy, double rx, double ry) {
    return static_cast<Sign>(del3d_orient2d(px, py, qx, qy, rx, ry));
}

Sign coplanar_orientation(const double* p, const double* q, const double* r) {
    return static_cast<Sign>(del3d_coplanar_orientation(p, q, r));
}

Sign coplanar_orientation(const double* p, const double* q, const double* r, const double* s) {
    return static_cast<Sign>(del3d_coplanar_orientation4(p, q, r, s));
}

bool collinear(const double* p, const double* q, const double* r) {
    return del3d_collinear(p, q, r) != 0;
}

bool equal(const double* p, const double* q) { return del3d_equal(p, q) != 0; }

CollinearPosition collinear_position(const double* p, const double* q, const double* r) {
    return static_cast<CollinearPosition>(del3d_collinear_position(p, q, r));
}

Sign coplanar_in_circle(const double* p0, const double* p1, const double* p2, const double* p) {
    return static_cast<Sign>(del3d_coplanar_in_circle(p0, p1, p2, p));
}

Sign coplanar_in_circle_unperturbed(const double* p0, const double* p1, const double* p2,
                                    const double* p) {
    return static_cast<Sign>(del3d_coplanar_in_circle_unperturbed(p0, p1, p2, p));
}

Point3 circumcenter(const Point3& p0, const Point3& p1, const Point3& p2, const Point3& p3) {
    Point3 r;
    del3d_circumcenter(&p0.x, &p1.x, &p2.x, &p3.x, 0, &r.x);
    return r;
}

Point3 circumcenter_lazy(const Point3& p, const Point3& q, const Point3& r, const Point3& s) {
    Point3 o;
    del3d_circumcenter(&p.x, &q.x, &r.x, &s.x, 1, &o.x);
    return o;
}

void spatial_sort(const std::vector<Point3>& points, std::vector<int>& order) {
    if (order.size() < 2) return;
    static_assert(sizeof(Point3) == 24, "Point3 must be three packed doubles");
    const int rc = del3d_spatial_sort_host(points.empty() ? nullptr : &points[0].x,
                                           static_cast<int64_t>(points.size()), order.data(),
                                           static_cast<int64_t>(order.size()));
    if (rc != DEL3D_OK) {
        std::fprintf(stderr, "del3d_amd: spatial_sort failed: %s\n", del3d_last_error());
        std::abort();
    }
}

}  // namespace del3d
