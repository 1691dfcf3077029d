// host_scalar.h - single queries evaluated on the host (host_scalar.cpp).
#pragma once

#include <cstdint>

#include "pred_core.h"

namespace d3 {

// orient3d / in_sphere / in_sphere_unperturbed (DEL3D_OP_*) of the points
// p[0..4] (orient3d reads four); id[k] = the slot's pointer identity (equal
// ids are equal pointers, src/predicates.cpp:220-227). Returns the sign.
int host_pred3(uint8_t op, const P3 p[5], const int32_t id[5]);

// A coplanar predicate (DEL3D_OP2_*) of p[0..3] in argument order, ids as
// above; returns the predicate's result as del3d_predicates2 writes it.
int host_pred2(uint8_t op, const P3 p[4], const int32_t id[4]);

// circumcenter (lazy = 0) or circumcenter_lazy of the tetrahedron p[0..3].
void host_circumcenter(const P3 p[4], int lazy, double out[3]);

}  // namespace d3
