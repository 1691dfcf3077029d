/*
 * oracle.h - CPU restatement of del3d's hot path, used ONLY as a checker.
 *
 * TEST INFRASTRUCTURE. Nothing in the product (del3d_amd/, include/) links,
 * loads or calls this library. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only to check or to time a CPU
 * baseline beside the GPU.
 *
 * What it restates (reference paths relative to the del3d source tree):
 *   - predicates   src/predicates.cpp:22-232 (filter, expansion exact stage,
 *                  symbolic perturbation) with src/expansion.h:33-178
 *   - circumcentres src/triangulation.cpp:805-930 with src/interval.h:27-186
 *   - coplanar     src/predicates2.cpp:22-259 (orient2d, coplanar orientation,
 *                  collinear, equal, collinear_position, in-circle) and the
 *                  driver's plain-double orientation, src/triangulation.cpp:296-308
 *   - an independent big-integer predicate oracle following the method of
 *                  tests/exact.h:161-291 (power-of-two frame, integer dets)
 *   - spatial sort src/spatial_sort.cpp:17-151 with src/rand48.h:26-153
 *   - point sets   tests/pointsets.h:36-126 (mt19937 + explicit mappings)
 *
 * Pinning (what checks each part against something it does not derive from):
 *   - raw predicate signs: the big-integer oracle (a different method) and
 *     the known answers of tests/test_predicates.cpp:64-76;
 *   - perturbation: the rule text of include/del3d/predicates.h:28-48 and a
 *     second restatement over the big-integer oracle;
 *   - coplanar: the known answers of tests/test_predicates2.cpp:64-90 and the
 *     big-integer oracle's apex method (tests/exact.h:294-378) on exactly
 *     coplanar queries; the compiled reference where it builds;
 *   - rand48 draws: glibc srand48/lrand48 (same LCG constants and output);
 *   - spatial sort: the 17 pinned FNV-1a hashes of
 *     tests/test_spatial_sort.cpp:115-147 and, where the reference builds,
 *     the compiled reference (oracle/_ref) on larger inputs.
 */
#ifndef DEL3D_ORACLE_H
#define DEL3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Predicate op codes shared with the batched API. */
enum { ORC_OP_ORIENT3D = 0, ORC_OP_INSPHERE = 1, ORC_OP_INSPHERE_RAW = 2 };

/* ---- filtered + expansion predicates (restating src/predicates.cpp) ---- */
int orc_orient3d(const double* a, const double* b, const double* c, const double* d);
int orc_in_sphere_unperturbed(const double* a, const double* b, const double* c,
                              const double* d, const double* e);
int orc_in_sphere(const double* a, const double* b, const double* c,
                  const double* d, const double* e);

/* Batched form over an AoS point array xyz[3*n_pts] and query records
 * idx[5*q .. 5*q+4] (orient3d uses the first four). Writes -1/0/+1. */
void orc_pred_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                    const int32_t* idx5, int64_t n_q, int8_t* sign_out);

/* Stage counters of the last orc_pred_batch call on this thread:
 * [0]=filter-decided, [1]=exact stage entered, [2]=perturbation entered. */
void orc_last_stage_counts(int64_t out[3]);

/* ---- coplanar predicates (restating src/predicates2.cpp) --------------- */
/* Op codes of the coplanar batch (idx[5*q ..]: the op's points in argument
 * order). Results: orientations -1/0/+1, collinear/equal 0/1,
 * collinear_position 0..4 (before, source, middle, target, after). */
enum {
    ORC_OP2_COPLANAR_ORIENTATION = 0,  /* (p,q,r) */
    ORC_OP2_COPLANAR_ORIENTATION4 = 1, /* (p,q,r,s) */
    ORC_OP2_COLLINEAR = 2,             /* (p,q,r) */
    ORC_OP2_EQUAL = 3,                 /* (p,q) */
    ORC_OP2_COLLINEAR_POSITION = 4,    /* (p,q,r) */
    ORC_OP2_IN_CIRCLE = 5,             /* (p0,p1,p2,p), perturbed */
    ORC_OP2_IN_CIRCLE_UNPERTURBED = 6, /* (p0,p1,p2,p) */
    ORC_OP2_ORIENT2D_XY = 7,           /* orient2d on the xy coordinates of (p,q,r) */
    ORC_OP2_INEXACT_ORIENTATION = 8    /* the driver's plain-double orientation (p,q,r,s) */
};
/* Points an op reads. */
int orc_pred2_points(uint8_t op);
int orc_pred2_one(uint8_t op, const double* a, const double* b, const double* c,
                  const double* d);
void orc_pred2_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                     const int32_t* idx5, int64_t n_q, int8_t* sign_out);

/* ---- circumcentres (restating src/triangulation.cpp:805-930) ----------- */
/* cells: 4 point indices per cell; out: 3 doubles per cell. lazy != 0:
 * circumcenter_lazy (intervals under FE_UPWARD, src/interval.h), else the
 * exact circumcenter. */
void orc_circumcenters(const double* xyz, const int32_t* cells, int64_t n, int lazy, double* out);

/* ---- big-integer oracle (method of tests/exact.h) ---------------------- */
int orc_int_orient3d(const double* a, const double* b, const double* c, const double* d);
int orc_int_in_sphere_raw(const double* a, const double* b, const double* c,
                          const double* d, const double* e);
int orc_int_in_sphere(const double* a, const double* b, const double* c,
                      const double* d, const double* e);
void orc_int_pred_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                        const int32_t* idx5, int64_t n_q, int8_t* sign_out);
/* Coplanar predicates by the method of tests/exact.h:294-378: orient2d on an
 * integer frame, and the in-circle test as an in-sphere test against the apex
 * p0 + (p1-p0) x (p2-p0). The in-circle answers equal the library's only for
 * exactly coplanar queries (its precondition). Same op codes as above; for
 * ORC_OP2_INEXACT_ORIENTATION it gives the exact orientation sign, which the
 * plain-double one must equal wherever its error bound is small. */
int orc_int_orient2d(double px, double py, double qx, double qy, double rx, double ry);
void orc_int_pred2_batch(const double* xyz, int64_t n_pts, const uint8_t* op,
                         const int32_t* idx5, int64_t n_q, int8_t* sign_out);

/* ---- spatial sort (restating src/spatial_sort.cpp + src/rand48.h) ------ */
/* order[0..n) is permuted in place; xyz is AoS (Point3). */
void orc_spatial_sort(const double* xyz, int64_t n, int32_t* order);
/* Only the rand48 Fisher-Yates shuffle of spatial_sort (no Hilbert pass). */
void orc_shuffle(int64_t n, int32_t* order);
/* Number of rejected draws the shuffle of n elements makes. */
int64_t orc_shuffle_rejections(int64_t n);
/* The raw rand48 output stream for a seed (31-bit values). */
void orc_rand48_draws(uint32_t seed, int64_t count, uint32_t* out);
/* Total Lomuto element-visits spatial_sort makes on this input. */
int64_t orc_spatial_sort_visits(const double* xyz, int64_t n);
/* FNV-1a 64 over the 4 low bytes of each index (tests/test_spatial_sort.cpp:46-51). */
uint64_t orc_hash_order(const int32_t* order, int64_t n);

/* ---- point sets (restating tests/pointsets.h) -------------------------- */
/* Each writes 3*count doubles into out (caller sizes it; the count is
 * returned and is computable up front: see the docstrings in pointsets.c). */
int64_t orc_pts_cylinder(int64_t n, int q, uint32_t seed, double* out);
int64_t orc_pts_sphere(int64_t n, int q, uint32_t seed, double* out);
int64_t orc_pts_torus(int64_t n, int q, uint32_t seed, double* out);
int64_t orc_pts_box_shell(int m, double* out);
int64_t orc_pts_lattice(int nx, int ny, int nz, double* out);
int64_t orc_pts_random_cube(int64_t n, uint32_t seed, double* out);
/* mt19937 words, for pinning the generator itself. */
void orc_mt19937_words(uint32_t seed, int64_t count, uint32_t* out);
/* Index draws of tests/test_predicates.cpp:118-121: Rng(seed).below(n) x count. */
void orc_rng_below(uint32_t seed, uint64_t n, int64_t count, int64_t* out);

#ifdef __cplusplus
}
#endif

#endif
