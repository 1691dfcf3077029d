// Entry points of the emulated build (test infrastructure): C wrappers over
// the library's C ABI, run on the CPU by hipemu.h.
#include <cstdint>
#include <cstdlib>

#include <hip/hip_runtime.h>

#include "../../include/del3d_amd.h"
#include "../../del3d_amd/csrc/predicates.h"

extern "C" int emu_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order) {
    return del3d_spatial_sort_host(xyz, n_pts, order, n_order);
}

extern "C" int emu_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                                     int32_t* order) {
    // device buffers are host memory in the emulation
    return del3d_spatial_sort_sets(xyz, off, n_sets, order, nullptr);
}

extern "C" int emu_predicates(const double* xyz, int64_t n_pts, const uint8_t* op,
                              const int32_t* idx, int64_t n_q, int8_t* out) {
    return del3d_predicates_host(xyz, n_pts, op, idx, n_q, out);
}

extern "C" int emu_predicates_stats(const double* xyz, int64_t n_pts, const uint8_t* op,
                                    const int32_t* idx, int64_t n_q, int8_t* out, int64_t* st) {
    return del3d_predicates_stats(xyz, n_pts, op, idx, n_q, out, nullptr, st);
}

extern "C" int emu_predicates_forced(const double* xyz, int64_t n_pts, const uint8_t* op,
                                     const int32_t* idx, int64_t n_q, int8_t* out, int path) {
    // path 1 = LDS kernel, 2 = global kernel: a one-set plan whose routing
    // threshold sends the set to the chosen kernel
    const int64_t po[2] = {0, n_pts}, qo[2] = {0, n_q};
    del3d_pred_plan p = nullptr;
    int rc = del3d_pred_plan_create(po, qo, 1, &p);
    if (rc) return rc;
    int32_t knobs[DEL3D_PLAN_KNOBS];
    del3d_pred_plan_get_knobs(p, knobs);
    knobs[0] = path == 1 ? 5632 : 0;
    knobs[1] = 2048;   // small tasks: several per set, so task switching is exercised
    knobs[2] = 512;
    rc = del3d_pred_plan_set_knobs(p, knobs, DEL3D_PLAN_KNOBS);
    int64_t st[6];
    if (!rc) rc = del3d_pred_plan_run_stats(p, xyz, op, idx, out, nullptr, st);
    del3d_pred_plan_destroy(p);
    return rc;
}

extern "C" int emu_predicates_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                                   const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                                   int8_t* out) {
    return del3d_predicates_sets(xyz, pt_off, n_sets, op, idx, q_off, out, nullptr);
}

// A multi-set plan run with the given global task size and XCD binning, all
// sets routed to the global kernel (task stealing between bins is exercised
// when a bin runs dry before the others).
extern "C" int emu_plan_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                             const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                             int8_t* out, int glob_q, int binned) {
    del3d_pred_plan p = nullptr;
    int rc = del3d_pred_plan_create(pt_off, q_off, n_sets, &p);
    if (rc) return rc;
    int32_t knobs[DEL3D_PLAN_KNOBS];
    del3d_pred_plan_get_knobs(p, knobs);
    knobs[0] = 0;
    knobs[2] = glob_q;
    knobs[3] = 1;
    knobs[4] = binned;
    rc = del3d_pred_plan_set_knobs(p, knobs, DEL3D_PLAN_KNOBS);
    int64_t st[6];
    if (!rc) rc = del3d_pred_plan_run_stats(p, xyz, op, idx, out, nullptr, st);
    del3d_pred_plan_destroy(p);
    return rc;
}

extern "C" int emu_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op,
                               const int32_t* idx, int64_t n_q, int8_t* out) {
    return del3d_predicates2_host(xyz, n_pts, op, idx, n_q, out);
}

extern "C" int emu_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n,
                                 int lazy, double* out) {
    return del3d_circumcenters_host(xyz, n_pts, cells, n, lazy, out);
}

extern "C" int emu_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op,
                                      const int32_t* idx, int64_t n_q, int8_t* out) {
    // device buffers are host memory in the emulation; the call is
    // stream-ordered, so wait for it (the deferred mode runs kernels only here)
    const int rc = del3d_predicates_noalloc(xyz, n_pts, op, idx, n_q, out, nullptr);
    hipStreamSynchronize(nullptr);
    return rc;
}

// The one-pass small-batch kernel (launch_predicates_small, the driver
// bridge's 3D waves) with its wide-tier tail.
extern "C" int emu_predicates_small(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                                    int64_t n_q, int8_t* out) {
    d3::PredCounters* ctr = nullptr;
    uint32_t* wide = nullptr;
    if (hipMalloc((void**)&ctr, sizeof(d3::PredCounters)) != hipSuccess) return 1;
    if (hipMalloc((void**)&wide, (size_t)n_q * 4 + 4) != hipSuccess) return 1;
    int rc = 0;
    try {
        d3::launch_predicates_small(xyz, n_pts, op, idx, n_q, out, ctr, wide, nullptr);
        if (hipDeviceSynchronize() != hipSuccess) rc = 1;
    } catch (...) {
        rc = 1;
    }
    (void)hipFree(ctr);
    (void)hipFree(wide);
    return rc;
}
