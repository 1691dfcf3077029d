// dropin.cpp - the reference's C++ symbols over the C ABI.
//
// Scalar predicate calls run one query on the GPU (del3d_orient3d & co.);
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
