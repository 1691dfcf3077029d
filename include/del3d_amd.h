/*
 * del3d_amd.h - C ABI of the MI355X (gfx950) implementation of del3d's hot path:
 * the exact filtered predicates orient3d / in_sphere (with symbolic
 * perturbation), their coplanar counterparts, and the rand48-shuffled
 * multiscale Hilbert spatial sort.
 *
 * Every result is bit-exact with the reference on the same input: predicate
 * signs are the signs the reference returns, sort permutations are the
 * permutations it produces.
 *
 * Conventions
 *   - Plain pointers and sizes; no framework types. "Device pointer" means
 *     memory the HIP runtime can address from the current device (hipMalloc,
 *     or a framework's device tensor).
 *   - `stream` is a hipStream_t passed as void*; NULL is the null stream.
 *     Calls are stream-ordered and return once the work is enqueued, except
 *     where stated. Scratch memory comes from a per-stream cache of device
 *     blocks, so concurrent calls on different streams are safe.
 *   - Return value: DEL3D_OK (0) or a DEL3D_ERR_* code; del3d_last_error()
 *     gives the message of the calling thread's last failure.
 *   - Points are the reference's Point3 (include/del3d/spatial_sort.h:31):
 *     x, y, z as three consecutive doubles, 24 bytes per point.
 *   - Inputs must be finite (include/del3d/predicates.h:25 assumes the same).
 */
#ifndef DEL3D_AMD_H
#define DEL3D_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEL3D_OK 0
#define DEL3D_ERR_INVALID 1     /* bad argument (size, pointer, op code) */
#define DEL3D_ERR_DEVICE 2      /* a HIP call failed */
#define DEL3D_ERR_NOMEM 3       /* device allocation failed */
#define DEL3D_ERR_INTERNAL 4    /* a device-side consistency check failed (no result written) */

/* Predicate op codes, one byte per query. */
#define DEL3D_OP_ORIENT3D 0              /* orient3d(r0, r1, r2, r3) */
#define DEL3D_OP_IN_SPHERE 1             /* in_sphere(r0, r1, r2, r3, r4) */
#define DEL3D_OP_IN_SPHERE_UNPERTURBED 2 /* in_sphere_unperturbed(r0..r4) */

/* Coplanar predicate op codes (del3d_predicates2), one byte per query. The
 * query's points are idx[5q+0 ..] in the reference's argument order. */
#define DEL3D_OP2_COPLANAR_ORIENTATION 0   /* coplanar_orientation(p,q,r): -1/0/+1 */
#define DEL3D_OP2_COPLANAR_ORIENTATION4 1  /* coplanar_orientation(p,q,r,s): -1/0/+1 */
#define DEL3D_OP2_COLLINEAR 2              /* collinear(p,q,r): 0/1 */
#define DEL3D_OP2_EQUAL 3                  /* equal(p,q): 0/1 */
#define DEL3D_OP2_COLLINEAR_POSITION 4     /* collinear_position(p,q,r): 0..4 */
#define DEL3D_OP2_IN_CIRCLE 5              /* coplanar_in_circle(p0,p1,p2,p): -1/+1 */
#define DEL3D_OP2_IN_CIRCLE_UNPERTURBED 6  /* coplanar_in_circle_unperturbed: -1/0/+1 */
#define DEL3D_OP2_ORIENT2D_XY 7            /* orient2d(p.x,p.y, q.x,q.y, r.x,r.y): -1/0/+1 */
#define DEL3D_OP2_INEXACT_ORIENTATION 8    /* the Delaunay driver's plain-double pre-walk
                                              orientation (p,q,r,s), src/triangulation.cpp:
                                              296-308: -1/0/+1, not exact by design */

/* Mixed batches (del3d_predicates_mixed, and the predicate trace format):
 * op bytes 0-2 are the DEL3D_OP_* codes, DEL3D_MIXED_OP2_BASE + k the
 * DEL3D_OP2_* code k. */
#define DEL3D_MIXED_OP2_BASE 16

/* Message of the calling thread's last failed call ("" if none). */
const char* del3d_last_error(void);

/* Library version string. */
const char* del3d_version(void);

/*
 * Device memory the library keeps between calls (no counterpart in the
 * reference): scratch blocks are cached per stream after a call (at most
 * 8 GiB per stream; a block beyond that is freed), so later calls on the
 * stream reuse them, and pinned staging buffers serve the host-memory forms.
 * A hipMalloc that runs out of memory empties the cache and is retried once.
 * del3d_release_cache waits for the current device and frees all of it (the
 * next call allocates again).
 */
int del3d_release_cache(void);

/*
 * Batched predicates over device-resident data.
 *
 * Replaces, per query q, one call of
 *   orient3d(a,b,c,d)               include/del3d/predicates.h:26 (src/predicates.cpp:125)
 *   in_sphere(a,b,c,d,e)            include/del3d/predicates.h:49-50 (src/predicates.cpp:206)
 *   in_sphere_unperturbed(a,..,e)   include/del3d/predicates.h:54-55 (src/predicates.cpp:149)
 * with a = xyz + 3*idx[5q+0], ..., e = xyz + 3*idx[5q+4]. Equal indices are
 * the reference's equal pointers (the perturbation compares pointer
 * identity, src/predicates.cpp:220-227); distinct indices of equal
 * coordinates are distinct pointers. idx[5q+4] is ignored for orient3d.
 *
 *   xyz      device, 3*n_pts doubles
 *   op       device, n_q bytes (DEL3D_OP_*)
 *   idx      device, 5*n_q int32, each in [0, n_pts)
 *   sign_out device, n_q int8, written with -1 / 0 / +1
 * Out-of-range pointers are undefined behaviour in the reference; here a
 * query with an index outside [0, n_pts) or an unknown op code gets sign 0
 * (counted in stats[5] of the stats form), and the device never reads
 * outside xyz.
 */
int del3d_predicates(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                     int64_t n_q, int8_t* sign_out, void* stream);

/*
 * As del3d_predicates, and additionally returns, after synchronizing the
 * stream, how the queries were decided (host array of 6 int64):
 *   [0] by the FP64 semi-static filter
 *   [1] by the certified small-integer re-evaluation in registers
 *   [2] by exact integer arithmetic in registers (1-4 limbs per difference)
 *   [3] by exact integer arithmetic in LDS (exponent spans above 2^119)
 *   [4] in_sphere queries that needed the symbolic perturbation
 *   [5] queries with a point index outside the point set or an unknown op
 *       code (sign 0 written; the reference's behaviour for such a pointer is
 *       undefined)
 */
int del3d_predicates_stats(const double* xyz, int64_t n_pts, const uint8_t* op,
                           const int32_t* idx, int64_t n_q, int8_t* sign_out, void* stream,
                           int64_t* stats_host);

/*
 * Measurement helper (no counterpart in the reference): as del3d_predicates,
 * and additionally records HIP events around each kernel on `stream`,
 * synchronizes it, and writes the kernels' durations in milliseconds:
 *   ms_out[0] fast path (filter, certified grid, in-register perturbation)
 *   ms_out[1] exact integer tier (registers)   ms_out[2] wide exact tier (LDS)
 */
int del3d_predicates_timed(const double* xyz, int64_t n_pts, const uint8_t* op,
                           const int32_t* idx, int64_t n_q, int8_t* sign_out, void* stream,
                           float* ms_out);

/*
 * Measurement helper (no counterpart in the reference): as del3d_predicates
 * (at most 2^28 queries: one launch sequence), and records the caller's events (hipEvent_t passed
 * as void*) on `stream` immediately before and after the fast-path kernel,
 * without synchronizing, so a timed loop can bracket every launch.
 */
int del3d_predicates_ev(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* sign_out, void* stream, void* ev_begin, void* ev_end);

/*
 * Allocation-free form of del3d_predicates (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same
 * arguments, results and stream semantics as del3d_predicates, but every
 * scratch buffer (the exact tiers' queues, counters, the set's grid summary
 * and its float or packed copy) lives in one file-scope device array of
 * 32 MiB that each device holds until the process ends; the call makes no
 * device allocation. Limits: n_pts <= 1,048,576 (larger sets are refused
 * with DEL3D_ERR_INVALID); any n_q (processed in launch sequences of 2^20).
 * Calls that overlap on one device share the array and race: serialize them
 * (one stream, or wait between calls). The array's contents are whatever the
 * last call left; the kernels clear what they need.
 */
int del3d_predicates_noalloc(const double* xyz, int64_t n_pts, const uint8_t* op,
                             const int32_t* idx, int64_t n_q, int8_t* sign_out, void* stream);

/*
 * Many point sets in one call (the batched configuration). Set s owns the
 * points xyz[3*pt_off[s] .. 3*pt_off[s+1]) and the queries
 * [q_off[s], q_off[s+1]) of op/idx/sign_out, whose indices are local to the
 * set (in [0, pt_off[s+1]-pt_off[s])). Each query is answered exactly as
 * del3d_predicates would answer it on its set alone.
 *   pt_off, q_off: host arrays of n_sets+1 int64, non-decreasing, from 0;
 *   at most 2^32-1 queries in total.
 * Builds a plan (below), runs it, synchronizes `stream`, frees the plan.
 */
int del3d_predicates_sets(const double* xyz, const int64_t* pt_off, int64_t n_sets,
                          const uint8_t* op, const int32_t* idx, const int64_t* q_off,
                          int8_t* sign_out, void* stream);

/*
 * Plans (no counterpart in the reference, whose predicates are scalar
 * calls): the task tables of a fixed set layout, built and uploaded once, so
 * that a run only enqueues kernels - no allocation, copy or host
 * synchronization - and can be captured into a HIP graph.
 *
 *   del3d_pred_plan_create    layout as in del3d_predicates_sets; allocates
 *                             the plan's device tables and queues on the
 *                             current device (besides the tables, 16 bytes
 *                             per point and 8 per query, up to 2^28 queries:
 *                             a run goes in launch sequences of at most
 *                             2^28 queries).
 *   del3d_pred_plan_run       enqueues the batch on `stream`.
 *   del3d_pred_plan_run_stats as run, then synchronizes and returns the
 *                             6 counters of del3d_predicates_stats.
 *   del3d_pred_plan_run_graph captures a run into a graph on first use (and
 *                             whenever the buffers or the stream change),
 *                             then replays it with one graph launch.
 *   del3d_pred_plan_autotune  times candidate knob settings on these buffers
 *                             (synchronizing), keeps the fastest; *best_ms
 *                             receives its time. Results never depend on the
 *                             knobs, only speed does.
 *   del3d_pred_plan_set_knobs / get_knobs
 *                             knobs[0] largest set (points) routed to the
 *                             LDS kernel (at most 5632; 0 = none),
 *                             [1] queries per LDS task, [2] queries per
 *                             global task, [3] global workgroups per CU,
 *                             [4] 1 = each XCD takes the global tasks of
 *                             its own share of the sets first (sets stay
 *                             in that XCD's L2), 0 = one shared list.
 *   del3d_pred_plan_destroy   frees the plan; call it once its runs finished.
 * A plan holds one set of queues: runs of one plan must not overlap (use one
 * stream per plan, or one plan per stream).
 */
#define DEL3D_PLAN_KNOBS 5
typedef struct del3d_pred_plan_s* del3d_pred_plan;
int del3d_pred_plan_create(const int64_t* pt_off, const int64_t* q_off, int64_t n_sets,
                           del3d_pred_plan* plan_out);
int del3d_pred_plan_run(del3d_pred_plan plan, const double* xyz, const uint8_t* op,
                        const int32_t* idx, int8_t* sign_out, void* stream);
int del3d_pred_plan_run_stats(del3d_pred_plan plan, const double* xyz, const uint8_t* op,
                              const int32_t* idx, int8_t* sign_out, void* stream,
                              int64_t* stats_host);
int del3d_pred_plan_run_graph(del3d_pred_plan plan, const double* xyz, const uint8_t* op,
                              const int32_t* idx, int8_t* sign_out, void* stream);
int del3d_pred_plan_autotune(del3d_pred_plan plan, const double* xyz, const uint8_t* op,
                             const int32_t* idx, int8_t* sign_out, void* stream, float* best_ms);
int del3d_pred_plan_set_knobs(del3d_pred_plan plan, const int32_t* knobs, int32_t n);
int del3d_pred_plan_get_knobs(del3d_pred_plan plan, int32_t* knobs);
int del3d_pred_plan_destroy(del3d_pred_plan plan);

/*
 * Serving loop (no counterpart in the reference): continuous batching of
 * predicate and sort jobs on resident point sets, under an HBM budget.
 *
 *   del3d_server_create      budget in bytes for resident sets plus batch
 *                            buffers; n_slots batches in flight at once (1..8,
 *                            each with its own stream); at most
 *                            max_batch_queries (<= 2^26) queries per batch and
 *                            per job. Starts the scheduler thread.
 *   del3d_server_put_points  uploads a point set (host array of 3*n_pts
 *                            doubles) and keeps it resident. The host array
 *                            must stay valid until the set is dropped: a set
 *                            no running batch uses may be evicted (least
 *                            recently used first) to admit other work and is
 *                            uploaded again from it when a job needs it.
 *   del3d_server_drop_points forgets a set (it must have no pending job).
 *   del3d_server_submit      queues a job: n_q queries (op, idx as in
 *                            del3d_predicates, host arrays, indices local to
 *                            the set); signs are written to the host array
 *                            sign_out. The arrays must stay valid until
 *                            del3d_server_wait returns for the job. Returns at
 *                            once; jobs that arrive while batches run are
 *                            admitted into the next free slot.
 *   del3d_server_submit_sort queues a sort job: the Hilbert spatial sort of the
 *                            whole resident set (order 0..n-1, as
 *                            del3d_spatial_sort sorts it); the permutation is
 *                            written to the host array order_out (n_pts int32),
 *                            which must stay valid until the job is waited for.
 *                            A sort job runs as a batch of its own; its scratch
 *                            (about 96 B per point + 64 MiB) counts against the
 *                            budget while it runs.
 *   del3d_server_wait        blocks until the job finished; latency_ms (may be
 *                            NULL) receives submit-to-completion time. Returns
 *                            the job's status (DEL3D_ERR_NOMEM if its set or
 *                            batch could not fit the budget).
 *   del3d_server_info        info[8]: device bytes used, resident set bytes,
 *                            resident sets, evictions, uploads, batches, jobs
 *                            finished, budget.
 *   del3d_server_destroy     lets queued jobs finish, then frees everything.
 * Each job's signs are exactly those of del3d_predicates on its set alone,
 * each sort job's permutation that of del3d_spatial_sort. Host arrays that
 * are pinned (hipHostMalloc) are copied by DMA directly; pageable ones go
 * through the serving loop's pinned staging.
 */
typedef struct del3d_server_s* del3d_server;
int del3d_server_create(int64_t hbm_budget_bytes, int32_t n_slots, int64_t max_batch_queries,
                        del3d_server* server_out);
int del3d_server_put_points(del3d_server server, const double* xyz, int64_t n_pts, int64_t* set_id);
int del3d_server_drop_points(del3d_server server, int64_t set_id);
int del3d_server_submit(del3d_server server, int64_t set_id, const uint8_t* op, const int32_t* idx,
                        int64_t n_q, int8_t* sign_out, int64_t* job_id);
int del3d_server_submit_sort(del3d_server server, int64_t set_id, int32_t* order_out,
                             int64_t* job_id);
int del3d_server_wait(del3d_server server, int64_t job_id, double* latency_ms);
int del3d_server_info(del3d_server server, int64_t* info);
int del3d_server_destroy(del3d_server server);

/*
 * Spatial sort of device-resident data. Replaces
 *   void spatial_sort(const std::vector<Point3>& points, std::vector<int>& order)
 *   include/del3d/spatial_sort.h:37 (src/spatial_sort.cpp:141-151)
 * `order` (device, n_order int32 indices into the n_pts points, normally
 * 0..n-1) is permuted in place into the reference's order; `xyz` (device,
 * 3*n_pts doubles) is not modified. Indices must lie in [0, n_pts).
 * Synchronizes `stream` several times per call (the recursion plan is driven
 * from the host).
 */
int del3d_spatial_sort(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order,
                       void* stream);

/*
 * Many independent spatial sorts in one call. Set s owns points
 * xyz[3*off[s] .. 3*off[s+1]) and order[off[s] .. off[s+1]), whose entries
 * are indices local to the set (in [0, off[s+1]-off[s])). Each set is sorted
 * exactly as del3d_spatial_sort would sort it alone.
 *   off: host array of n_sets+1 int64, non-decreasing, off[0] = 0.
 */
int del3d_spatial_sort_sets(const double* xyz, const int64_t* off, int64_t n_sets,
                            int32_t* order, void* stream);

/*
 * Batched coplanar predicates over device-resident data.
 *
 * Replaces, per query q, one call of the reference's coplanar predicates
 * (include/del3d/predicates2.h:26-69, src/predicates2.cpp:111-259):
 *   orient2d                          predicates2.h:26  (src/predicates2.cpp:111)
 *   coplanar_orientation(p,q,r)       predicates2.h:33  (src/predicates2.cpp:122)
 *   coplanar_orientation(p,q,r,s)     predicates2.h:39-40 (src/predicates2.cpp:133)
 *   collinear                         predicates2.h:44  (src/predicates2.cpp:148)
 *   equal                             predicates2.h:47  (src/predicates2.cpp:158)
 *   collinear_position                predicates2.h:51-52 (src/predicates2.cpp:162)
 *   coplanar_in_circle                predicates2.h:64-65 (src/predicates2.cpp:232)
 *   coplanar_in_circle_unperturbed    predicates2.h:68-69 (src/predicates2.cpp:182)
 * with the op's points at xyz + 3*idx[5q+k], k = 0 .. (points of the op)-1;
 * the remaining idx entries are ignored. Equal indices are equal pointers
 * (the in-circle perturbation compares pointer identity,
 * src/predicates2.cpp:249-255). orient2d reads the x and y coordinates of
 * its three points. out[q] is the predicate's result as listed with the op
 * codes above (collinear_position: 0 before, 1 source, 2 middle, 3 target,
 * 4 after, the reference's CollinearPosition order). A query with an index
 * outside [0, n_pts) or an unknown op code gets 0.
 *   xyz device, 3*n_pts doubles; op device, n_q bytes (DEL3D_OP2_*);
 *   idx device, 5*n_q int32; out device, n_q int8.
 */
int del3d_predicates2(const double* xyz, int64_t n_pts, const uint8_t* op, const int32_t* idx,
                      int64_t n_q, int8_t* out, void* stream);

/*
 * A mixed batch: any of the 3D and coplanar predicates and the driver's
 * inexact orientation in one call, op bytes as above (DEL3D_OP_* or
 * DEL3D_MIXED_OP2_BASE + DEL3D_OP2_*), each query's indices and result as in
 * del3d_predicates / del3d_predicates2. This is the replay form of the
 * predicate trace format (the op, result and index records of a run of the
 * reference's Delaunay driver). Unknown op bytes give 0.
 */
int del3d_predicates_mixed(const double* xyz, int64_t n_pts, const uint8_t* op,
                           const int32_t* idx, int64_t n_q, int8_t* out, void* stream);

/*
 * Conflict regions of an incremental Delaunay insertion, with the conflict
 * tests batched per breadth-first wave: the driver bridge. Replaces the body
 * of the reference driver's flood fill
 *   Builder::find_conflicts(d, p, cells_out, facets_out)
 *                              src/triangulation.cpp:643-689
 * and its per-cell tests in_conflict / in_conflict_2 (:588-629). The region
 * is found in waves: every untested neighbour of the cells found in conflict
 * in the previous wave is tested, the tests of one wave over all regions of
 * the call in one mixed predicate batch; then the reference's depth-first
 * order is replayed over the answers, so the cells and facets come out
 * exactly as the reference's fill lists them (its star rebuild depends on
 * that order). The bridge reads the caller's cell tables in place and writes
 * nothing to them: the caller sets the marks the rebuild expects (the
 * region's cells InConflict; the neighbour across each listed facet
 * OnBoundary).
 *
 * A view of one triangulation data structure (strided int32 tables, so the
 * reference's Cell / Vertex arrays are read where they lie):
 *   cell_v[c*cell_stride + k]  vertex k of cell c; cell_n[c*cell_stride + k]
 *                              the neighbour opposite it (k < 4; dimension 2
 *                              uses k < 3); cells in [0, n_cells)
 *   vertex_point[u*vertex_stride] + point_base
 *                              the point index of vertex u, in [0, n_pts) of
 *                              the bridge's points; vertices in [0, n_vertices)
 *   infinite                   the infinite vertex; dimension 2 or 3
 * Many triangulations can share one bridge: concatenate their points and give
 * each view its point_base.
 *
 *   del3d_bridge_create     over n_pts host points (3 doubles each, kept by
 *                           pointer: the array must outlive the bridge and
 *                           must not change while it lives, since it is
 *                           uploaded to the current device once, at the
 *                           first GPU wave). A wave of fewer than gpu_min_queries tests
 *                           is answered on the host with the kernels'
 *                           per-query code (0 = every wave on the GPU;
 *                           INT64_MAX = host only, which needs no device).
 *   del3d_bridge_conflicts  n_regions regions at once: region r starts at
 *                           start_cells[r] (the located cell, in conflict by
 *                           construction, not tested) for the point index
 *                           points[r]. Its cells go to
 *                           cells_out[cells_off[r] .. cells_off[r+1]) and its
 *                           facets, as (cell, local index) int32 pairs, to
 *                           facets_out[2*facets_off[r] .. 2*facets_off[r+1]),
 *                           both in the reference's order; cells_off and
 *                           facets_off hold n_regions+1 entries. If a buffer
 *                           is too small the call returns DEL3D_ERR_INVALID
 *                           with the needed totals in cells_off[n_regions]
 *                           and facets_off[n_regions]. A bad index in a view
 *                           returns DEL3D_ERR_INVALID.
 *   del3d_bridge_info       info[6]: waves, GPU batches, conflict tests,
 *                           tests answered on the GPU, on the host, and
 *                           follow-up tests (a hull orientation of 0 decided
 *                           in the facet's plane).
 *   del3d_bridge_destroy    frees the bridge.
 * Calls on one bridge are serialized.
 */
typedef struct del3d_tds_view {
    const int32_t* cell_v;
    const int32_t* cell_n;
    int64_t cell_stride;
    int64_t n_cells;
    const int32_t* vertex_point;
    int64_t vertex_stride;
    int64_t n_vertices;
    int64_t point_base;
    int32_t infinite;
    int32_t dimension;
} del3d_tds_view;
typedef struct del3d_bridge_s* del3d_bridge;
int del3d_bridge_create(const double* xyz, int64_t n_pts, int64_t gpu_min_queries,
                        del3d_bridge* bridge_out);
int del3d_bridge_conflicts(del3d_bridge bridge, int32_t n_regions, const del3d_tds_view* views,
                           const int32_t* start_cells, const int32_t* points, int32_t* cells_out,
                           int64_t cells_cap, int64_t* cells_off, int32_t* facets_out,
                           int64_t facets_cap, int64_t* facets_off);
int del3d_bridge_info(del3d_bridge bridge, int64_t* info);
int del3d_bridge_destroy(del3d_bridge bridge);

/*
 * Batched circumcentres of tetrahedra over device-resident data.
 *
 * Replaces, per cell c, one call of
 *   circumcenter(p0,p1,p2,p3)        include/del3d/delaunay.h:29-30
 *                                    (src/triangulation.cpp:805)      lazy = 0
 *   circumcenter_lazy(p,q,r,s)       include/del3d/delaunay.h:48-49
 *                                    (src/triangulation.cpp:874)      lazy != 0
 * with the points at xyz + 3*cells[4c+k]; out[3c .. 3c+2] receives the
 * result's x, y, z, bit for bit the reference's (the exact form's expansion
 * components and the lazy form's upward-rounded intervals are reproduced; see
 * the range note in circ_core.h). A cell with an index outside [0, n_pts)
 * gets NaN coordinates.
 *   xyz device, 3*n_pts doubles; cells device, 4*n_cells int32;
 *   out device, 3*n_cells doubles.
 */
int del3d_circumcenters(const double* xyz, int64_t n_pts, const int32_t* cells, int64_t n_cells,
                        int lazy, double* out, void* stream);

/* Host-memory forms: copy in, compute on the current device, copy out,
 * synchronously. For callers without device buffers. */
int del3d_predicates_host(const double* xyz, int64_t n_pts, const uint8_t* op,
                          const int32_t* idx, int64_t n_q, int8_t* sign_out);
int del3d_predicates2_host(const double* xyz, int64_t n_pts, const uint8_t* op,
                           const int32_t* idx, int64_t n_q, int8_t* out);
int del3d_circumcenters_host(const double* xyz, int64_t n_pts, const int32_t* cells,
                             int64_t n_cells, int lazy, double* out);
int del3d_spatial_sort_host(const double* xyz, int64_t n_pts, int32_t* order, int64_t n_order);

/*
 * Scalar predicates for the C++ drop-in (see del3d_amd_dropin.hpp): one
 * query evaluated on the host with the kernels' own per-query code (the same
 * filter, exact integer tiers and perturbation, compiled for the CPU with the
 * same floating-point rules), since a device round trip would cost far more
 * than the query. They return the sign; on an internal failure they print
 * del3d_last_error() and abort, since the reference's signatures have no
 * error channel.
 */
int del3d_orient3d(const double* a, const double* b, const double* c, const double* d);
int del3d_in_sphere(const double* a, const double* b, const double* c, const double* d,
                    const double* e);
int del3d_in_sphere_unperturbed(const double* a, const double* b, const double* c,
                                const double* d, const double* e);

/* The coplanar predicates, scalar (include/del3d/predicates2.h:26-69); same
 * conventions. orient2d takes the six coordinates as the reference does. */
int del3d_orient2d(double px, double py, double qx, double qy, double rx, double ry);
int del3d_coplanar_orientation(const double* p, const double* q, const double* r);
int del3d_coplanar_orientation4(const double* p, const double* q, const double* r,
                                const double* s);
int del3d_collinear(const double* p, const double* q, const double* r);
int del3d_equal(const double* p, const double* q);
int del3d_collinear_position(const double* p, const double* q, const double* r);
int del3d_coplanar_in_circle(const double* p0, const double* p1, const double* p2,
                             const double* p);
int del3d_coplanar_in_circle_unperturbed(const double* p0, const double* p1, const double* p2,
                                         const double* p);

/* One circumcentre (lazy or exact), computed on the host like the scalar
 * predicates; out: x, y, z. */
void del3d_circumcenter(const double* p0, const double* p1, const double* p2, const double* p3,
                        int lazy, double* out);

#ifdef __cplusplus
}
#endif

#endif /* DEL3D_AMD_H */
