// del3d_amd_dropin.hpp - the reference's C++ entry points, served by the GPU.
//
// Declares, with the reference's exact signatures (so that code compiled
// against the reference's headers links against libdel3d_amd.so unchanged):
//   Sign orient3d(...)               include/del3d/predicates.h:26
//   Sign in_sphere(...)              include/del3d/predicates.h:49-50
//   Sign in_sphere_unperturbed(...)  include/del3d/predicates.h:54-55
//   void spatial_sort(points, order) include/del3d/spatial_sort.h:37
//   the coplanar predicates          include/del3d/predicates2.h:26-69
//     (orient2d, coplanar_orientation x2, collinear, equal,
//      collinear_position, coplanar_in_circle(_unperturbed))
//   Point3 circumcenter(...)         include/del3d/delaunay.h:29-30
//   Point3 circumcenter_lazy(...)    include/del3d/delaunay.h:48-49
// The enum and struct layouts are the reference's (predicates.h:21,
// predicates2.h:50, spatial_sort.h:31), which the mangled symbols and the
// data layout depend on.
#ifndef DEL3D_AMD_DROPIN_HPP
#define DEL3D_AMD_DROPIN_HPP

#include <vector>

namespace del3d {

enum Sign { NEGATIVE = -1, ZERO = 0, POSITIVE = 1 };

struct Point3 {
    double x, y, z;
};

Sign orient3d(const double* a, const double* b, const double* c, const double* d);
Sign in_sphere(const double* a, const double* b, const double* c, const double* d, const double* e);
Sign in_sphere_unperturbed(const double* a, const double* b, const double* c, const double* d,
                           const double* e);
void spatial_sort(const std::vector<Point3>& points, std::vector<int>& order);

enum CollinearPosition { CP_BEFORE, CP_SOURCE, CP_MIDDLE, CP_TARGET, CP_AFTER };

Sign orient2d(double px, double py, double qx, double qy, double rx, double ry);
Sign coplanar_orientation(const double* p, const double* q, const double* r);
Sign coplanar_orientation(const double* p, const double* q, const double* r, const double* s);
bool collinear(const double* p, const double* q, const double* r);
bool equal(const double* p, const double* q);
CollinearPosition collinear_position(const double* p, const double* q, const double* r);
Sign coplanar_in_circle(const double* p0, const double* p1, const double* p2, const double* p);
Sign coplanar_in_circle_unperturbed(const double* p0, const double* p1, const double* p2,
                                    const double* p);

Point3 circumcenter(const Point3& p0, const Point3& p1, const Point3& p2, const Point3& p3);
Point3 circumcenter_lazy(const Point3& p, const Point3& q, const Point3& r, const Point3& s);

}  // namespace del3d

#endif
