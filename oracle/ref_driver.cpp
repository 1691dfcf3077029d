// ref_driver.cpp - the reference Delaunay driver's plain-double orientation,
// compiled from the driver's own source where it lies.
//
// TEST INFRASTRUCTURE ONLY. The pre-walk orientation of src/triangulation.cpp
// (:296-308) is a private static member of a class internal to that file, so
// it has no symbol to call. This unit includes the unchanged source file
// (oracle/Makefile passes the reference tree on the include path) with its
// class's members made accessible, and exports a batch entry point over it.
// The result goes to oracle/_ref/libdel3d_ref_driver.so, a library of its own
// (this unit holds a second copy of the driver, so it is not linked with
// triangulation.o).
//
// Every standard header the driver and its headers use is included first, so
// that the access redefinition below reaches the reference's code only.
#include <algorithm>
#include <array>
#include <cassert>
#include <cfenv>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#define private public
#include "src/triangulation.cpp"
#undef private

extern "C" void ref_inexact_orientation_batch(const double* xyz, const int32_t* idx5, int64_t n_q,
                                              int8_t* out) {
    for (int64_t q = 0; q < n_q; ++q) {
        const int32_t* r = idx5 + 5 * q;
        out[q] = static_cast<int8_t>(del3d::Builder::inexact_orientation(
            xyz + 3 * int64_t(r[0]), xyz + 3 * int64_t(r[1]), xyz + 3 * int64_t(r[2]),
            xyz + 3 * int64_t(r[3])));
    }
}
