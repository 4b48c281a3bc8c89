/*
 * quijibo-mi355x C ABI: the drop-in boundary for QUIJIBO's rational-quaternion
 * Julia field path on MI355X (gfx950).
 *
 * Plain C: pointers, sizes and status codes only. Device pointers are HIP
 * device addresses; `stream` is a hipStream_t (NULL = the null stream). Every
 * call that takes a stream only enqueues work on it unless stated otherwise.
 * Functions return QJ_OK or an error code; qj_last_error() gives the message.
 * The reference reports numeric trouble as inf/NaN in the field and calls
 * exit(0) on fatal conditions; here fatal conditions are status codes.
 *
 * Reference interfaces replaced (theodorekim/QUIJIBO):
 *   OPTIMIZE_3D::computeLogScaledPowerRationalMap    src/solvers/OPTIMIZE_3D.h:161, .cpp:1272-1370
 *   OPTIMIZE_3D::computeRationalMapAP                src/solvers/OPTIMIZE_3D.cpp:1374-1568
 *   OPTIMIZE_3D::computeLogScaledPowerRationalMapCached  src/solvers/OPTIMIZE_3D.cpp:1572-1718
 *   the dormant GPU seam setTopRoots/setConsts/runMap  src/solvers/OPTIMIZE_3D.cpp:3842-3858
 *   TRIANGLE_MESH(const FIELD_3D&) / computeStagedMarchingCubes  src/geometry/TRIANGLE_MESH.h:42, .cpp:989-1114
 *   TRIANGLE_MESH(const FIELD_3D&, top, bottom, ...) / computeNonlinearMarchingCubes  TRIANGLE_MESH.h:45, .cpp:62-79,152
 *   OPTIMIZE_3D::computeScoreAP, computeExpGradientAP, compute{Top,Bottom}GradientAP,
 *     compute{Top,Bottom}BulkScalarGradientAP       src/solvers/OPTIMIZE_3D.cpp:3563, 3589, 8905, 10012, 9250, 10360
 *   OPTIMIZE_3D::computeLogRationalCenteredGradient (root-position gradient) src/solvers/OPTIMIZE_3D.cpp:2872-2951
 *   FIELD_3D::clampInfsToMean / clampNansToMean / setBorders  src/geometry/FIELD_3D.cpp:5182-5234, 5357-5382
 *   FIELD_3D::fieldMin / fieldMax / invertedDistance / invertedDistanceNormalized
 *                                                    src/geometry/FIELD_3D.cpp:2960-3032, 5241-5352
 *   OPTIMIZE_3D::weightScore / inverseDistanceScore / computeMaximumScaledPowerRationalScore
 *                                                    src/solvers/OPTIMIZE_3D.cpp:4484-4563
 */
#ifndef QJ_C_API_H
#define QJ_C_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t qj_status;
enum {
  QJ_OK = 0,
  QJ_ERR_INVALID = 1,     /* bad argument (sizes, null pointers, unsupported combination) */
  QJ_ERR_HIP = 2,         /* a HIP runtime call failed */
  QJ_ERR_NOMEM = 3,       /* device allocation failed */
  QJ_ERR_CAPACITY = 4,    /* request larger than the configured capacity */
  QJ_ERR_NUMERIC = 5,     /* inconsistent numeric state (the reference would exit(0)) */
  QJ_ERR_SHUTDOWN = 6,    /* server is stopping */
  QJ_ERR_BUSY = 7         /* request refused: the request loop's pending work is over its limit */
};

/* A FIELD_3D geometry (FIELD_3D.h:58,73-80; FIELD_3D.cpp:28-57): cell-centred,
 * x fastest, data[z*xRes*yRes + y*xRes + x]. rotation is a quaternion
 * (w,x,y,z) applied about center (FIELD_3D.cpp:818-841); identity (1,0,0,0). */
typedef struct qj_grid {
  int32_t res[3];
  double center[3];
  double lengths[3];
  double rotation[4];
} qj_grid;

/* A POLYNOMIAL_4D in factored form (POLYNOMIAL_4D.h:137-150): roots (n x 4,
 * slot order w,x,y,z), per-root powers and the uniform power scalar. Host
 * pointers. Term i is (q - root_i)^(power_scalar * powers[i]). */
typedef struct qj_poly {
  int32_t n_roots;
  const double* roots;
  const double* powers;
  double power_scalar;
} qj_poly;

/* map flags */
#define QJ_MAP_BLACK_HOLE 1u  /* stop iterating when |q| < 10*FLT_MIN (OPTIMIZE_3D.cpp:1358-1359) */
#define QJ_MAP_SKIP_ORIGIN 2u /* leave cells with |q0| < 1e-8 unwritten (OPTIMIZE_3D.cpp:1428, 1619) */
#define QJ_MAP_ADAPTIVE 4u    /* flag overflowed cells and recompute them with extended range
                                 (computeRationalMapAP's long double re-pass, OPTIMIZE_3D.cpp:1477-1567) */
#define QJ_MAP_LAGRANGE 8u    /* RATIONAL_4D::evaluateLagrange order (RATIONAL_4D.cpp:27-88): top and bottom
                                 factors interleaved, r = (r t_i) / b_i, then the leftovers; the map of the
                                 Lagrange mesher (TRIANGLE_MESH.cpp:106-123, 281-324). QJ_FP64 only; not
                                 with QJ_MAP_SKIP_ORIGIN or QJ_MAP_ADAPTIVE */
#define QJ_MAP_EXTENDED_RANGE 16u /* the range of the reference's default build (Real = long double,
                                 SETTINGS.h:10): factored products, quotient and |q| carry a separate binary
                                 exponent, so products of many roots stay finite where FP64 gives inf, NaN or
                                 -inf; FP64 mantissas, so values agree with a long double evaluation to the
                                 FP64 tolerance. Every precision, the mesher and the request loop; not with
                                 QJ_MAP_ADAPTIVE (which has its own re-pass) or QJ_MAP_LAGRANGE */

/* Map parameters: the OPTIMIZE_3D members the map reads (OPTIMIZE_3D.h:67,73,81)
 * plus the escape radius each reference variant hard-codes (2000 at
 * OPTIMIZE_3D.cpp:1291, 20 at :1388 and :1589). */
typedef struct qj_map_params {
  int32_t max_iter;
  double escape;
  double exp_scaling;
  double slice;
  uint32_t flags;
} qj_map_params;

/* Reference presets for qj_map_params.flags/escape */
#define QJ_PRESET_SCALED_POWER_ESCAPE 2000.0 /* computeLogScaledPowerRationalMap, flags = BLACK_HOLE */
#define QJ_PRESET_AP_ESCAPE 20.0             /* computeRationalMapAP, flags = SKIP_ORIGIN|ADAPTIVE */

typedef enum qj_precision {
  QJ_FP64 = 0,       /* every voxel in FP64, reference expression order */
  QJ_FP32_BAND = 1,  /* FP32 bulk; FP64 re-evaluation of the band |F| < band, of
                        non-finite results and of voxels whose escape or black-hole
                        test was within FP32's error of its threshold */
  QJ_FP64_EXACT = 2  /* every voxel in FP64, exponents exactly 1 and 2 as the exact
                        products b and b*b (a few ulps from the polar form, with its
                        special cases), other exponents polar; the FP32 path's band
                        arithmetic. qj_field_eval, qj_field_eval_noalloc and
                        qj_points_eval; not with QJ_MAP_LAGRANGE */
} qj_precision;

/* Counters accumulated by the evaluation calls on a workspace since the last
 * qj_workspace_reset_stats. */
typedef struct qj_eval_stats {
  int64_t voxels;            /* voxels evaluated (bulk pass) */
  int64_t iterations;        /* executed voxel-iterations in the bulk pass */
  int64_t band_voxels;       /* voxels re-evaluated in FP64 (QJ_FP32_BAND) */
  int64_t band_iterations;   /* executed voxel-iterations of the FP64 re-evaluation */
  int64_t adaptive_voxels;   /* cells flagged as overflowed (QJ_MAP_ADAPTIVE) */
  int64_t nonfinite;         /* non-finite results written */
  /* the sampled FP64 cross-check (qj_workspace_check), summed over fields */
  int64_t checked;           /* cells re-evaluated with the reference method in FP64 */
  int64_t check_sign_flips;  /* of those, cells where (float)F > 0 differs from FP64's */
  int64_t check_class_mismatch; /* cells whose class (finite, NaN, +inf, -inf) differs */
  double check_max_abs_dF;   /* largest |F - F_fp64| over checked cells finite in both */
  double check_max_rel_dF;   /* largest |F - F_fp64| / max(1, |F_fp64|) */
} qj_eval_stats;

typedef struct qj_rational qj_rational;   /* top/bottom root sets resident in HBM */
typedef struct qj_workspace qj_workspace; /* device scratch + counters for one stream */

/* ---- library ---- */
const char* qj_version(void);
const char* qj_last_error(void); /* thread-local message of the last failing call */
const char* qj_status_string(qj_status s);
qj_status qj_init(int32_t device); /* selects the device, loads kernels; idempotent */

/* ---- root sets: POLYNOMIAL_4D pair on device (replaces setTopRoots,
 *      OPTIMIZE_3D.cpp:3842-3850) ---- */
qj_status qj_rational_create(const qj_poly* top, const qj_poly* bottom, qj_rational** out);
/* Replace roots/powers in place (same root counts), enqueued on `stream`. */
qj_status qj_rational_update(qj_rational* f, const qj_poly* top, const qj_poly* bottom, void* stream);
qj_status qj_rational_counts(const qj_rational* f, int32_t* n_top, int32_t* n_bottom);
qj_status qj_rational_destroy(qj_rational* f);

/* ---- workspaces ---- */
qj_status qj_workspace_create(qj_workspace** out);
qj_status qj_workspace_destroy(qj_workspace* ws);
/* Synchronizes `stream`, then copies the counters. */
qj_status qj_workspace_stats(qj_workspace* ws, void* stream, qj_eval_stats* out);
qj_status qj_workspace_reset_stats(qj_workspace* ws, void* stream);

/* The error statement of every field (no counterpart in the reference, which
 * has one precision): with samples > 0, each qj_field_eval on ws re-evaluates
 * `samples` evenly spaced cells (cell (2i+1) N / (2 samples), i < samples)
 * with the reference method in FP64 (polar powers, reference expression
 * order; the extended-range re-pass for cells QJ_MAP_ADAPTIVE flags) after
 * the field is written, and adds the comparison to the workspace statistics
 * (qj_eval_stats.checked ...). One extra kernel per field, about the latency
 * of one FP64 evaluation per cell (0.5 ms for 32+32 roots, maxIter 4).
 * 0 (the default) turns it off; at most 2^24. */
qj_status qj_workspace_check(qj_workspace* ws, int64_t samples);

/* Kernel timing on a workspace (profiling hook, like the reference's TIMER,
 * TIMER.h:37-79): when enabled, qj_field_eval / qj_points_eval record HIP
 * events around the bulk kernel and the FP64 re-evaluation kernel on the
 * call's stream; qj_workspace_kernel_ms synchronizes and returns the summed
 * milliseconds and launch count since enabling. */
qj_status qj_workspace_timing(qj_workspace* ws, int32_t enable);
qj_status qj_workspace_kernel_ms(qj_workspace* ws, double* bulk_ms, double* band_ms, int64_t* launches);

/* ---- field evaluation ---- */

/* computeLogScaledPowerRationalMap (OPTIMIZE_3D.cpp:1272-1370) and its AP /
 * cached variants through flags: for every cell, q0 = (X, Y, Z, slice) at the
 * cell centre; while |q| < escape and it < max_iter: q = c * T(q) * B(q)^-1
 * (q = c * T(q) with no bottom roots); field = log|q|. d_field: device,
 * res[0]*res[1]*res[2] doubles, x fastest. band is used by QJ_FP32_BAND (the
 * reference-facing default is 0.05). */
/* qj_field_eval and qj_points_eval enqueue kernels only (no memset nodes, no
 * host synchronization), so they may be captured in a hipGraph once a first
 * call of at least the same size (and with QJ_MAP_ADAPTIVE if the captured
 * call has it) has sized the workspace's lists; a larger call reallocates.
 * The band list holds max(n/4, 65536) of the n cells (at most n); cells that
 * need FP64 beyond it are marked in d_field and evaluated by a sweep kernel,
 * so every band width gives the same field. */
qj_status qj_field_eval(qj_workspace* ws, const qj_grid* grid, const qj_rational* f,
                        const qj_map_params* params, qj_precision precision, double band,
                        double* d_field, void* stream);

/* qj_field_eval without a workspace and without any device allocation, for
 * callers that cannot allocate (code under hipGraph capture, tight serving
 * loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * Same arguments and the same field, bit for bit, as qj_field_eval. The band
 * and overflow lists live in a file-scope device array of 32 MiB per GPU
 * (held until the process ends) and the grid is evaluated in z-slabs of at
 * most 4194304 cells, so any grid whose x*y slice has at most 4194304 cells is
 * accepted (QJ_ERR_CAPACITY otherwise). Enqueues kernels only; no statistics
 * or timing. Every call on one GPU shares that array: calls that overlap on
 * one GPU (other streams, other threads) race and must be serialized by the
 * caller. qj_init sets the array up; otherwise the first call on a GPU does. */
qj_status qj_field_eval_noalloc(const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                                qj_precision precision, double band, double* d_field, void* stream);

/* Same map at n arbitrary spatial points (X, Y, Z) (device, n x 3 doubles). */
qj_status qj_points_eval(qj_workspace* ws, const double* d_points, int64_t n, const qj_rational* f,
                         const qj_map_params* params, qj_precision precision, double band,
                         double* d_out, void* stream);

/* ---- marching cubes (TRIANGLE_MESH(const FIELD_3D&), TRIANGLE_MESH.cpp:48-57,
 *      computeStagedMarchingCubes :989-1114) ----
 * Flag bit k of a cell is ((float)(v_k - iso) > 0), corners ordered
 * NNN,NNP,NPN,NPP,PNN,PNP,PPN,PPP. One vertex per edge referenced by a cell's
 * case row, numbered by (lower endpoint linear index, axis x<y<z); vertex =
 * p0 + d0/(d0-d1) * (p1-p0) at cell centres; triangles in cell order then
 * table order; triangles with an edge shorter than 1e-7 are dropped.
 * iso = 0 is the reference behaviour (it has no isovalue).
 * qj_mc_run synchronizes `stream` and returns the counts; the mesh stays in
 * the workspace until the next qj_mc_run on it. */
qj_status qj_mc_run(qj_workspace* ws, const double* d_field, const qj_grid* grid, double iso,
                    int64_t* n_vertices, int64_t* n_triangles, void* stream);
/* Marching cubes of n fields (tiles) in one call, with every count kept on
 * the device and read back once at the end (no host synchronisation per
 * field). Each field is meshed as qj_mc_run would mesh it alone; the meshes
 * are stored one after another in the workspace: field k's vertices follow
 * those of fields 0..k-1, and so do its triangles, whose vertex indices are
 * local to field k. n_vertices / n_triangles (host, n entries each) receive
 * the counts. qj_mc_copy / qj_mc_device_ptrs give the concatenated mesh.
 * Synchronizes `stream`. */
qj_status qj_mc_run_tiles(qj_workspace* ws, int32_t n, const double* const* d_fields, const qj_grid* grids,
                          double iso, int64_t* n_vertices, int64_t* n_triangles, void* stream);
/* qj_mc_run_tiles writing the meshes straight into caller device buffers (no
 * copy out of the workspace): d_vertices holds vertex_capacity x 3 doubles,
 * d_triangles triangle_capacity x 3 int32, in the same order. When the meshes
 * need more, returns QJ_ERR_CAPACITY with the counts filled in (the sizes to
 * pass next time) and the buffers' contents undefined. The workspace keeps no
 * mesh: qj_mc_copy and the edge cache do not apply. Synchronizes `stream`.
 * No counterpart in the reference (its meshes live in host memory). */
qj_status qj_mc_run_tiles_into(qj_workspace* ws, int32_t n, const double* const* d_fields, const qj_grid* grids,
                               double iso, double* d_vertices, int64_t vertex_capacity, int32_t* d_triangles,
                               int64_t triangle_capacity, int64_t* n_vertices, int64_t* n_triangles,
                               void* stream);
/* Copy the last mesh: vertices n_vertices x 3 doubles, triangles n_triangles x 3
 * int32 (0-based). Pointers may be host or device memory (hipMemcpyDefault). */
qj_status qj_mc_copy(qj_workspace* ws, double* vertices, int32_t* triangles, void* stream);
/* Device pointers to the last mesh (valid until the next qj_mc_run on ws). */
qj_status qj_mc_device_ptrs(qj_workspace* ws, const double** d_vertices, const int32_t** d_triangles);

/* Non-linear low-memory marching cubes (the TRIANGLE_MESH(center, lengths, res,
 * top, bottom, expScaling, maxIterations, slice, isosurface, rotation[, cache])
 * mesher, TRIANGLE_MESH.cpp:84-101, 681-768): the map is evaluated slab by slab
 * (chunk_slices z-slices at a time plus a 3-slice halo; never the whole field),
 * cases come from log|q| (TRIANGLE_MESH.cpp:379), each vertex is found by
 * bisection on log|q| - iso re-evaluated in FP64 (at most 6 halvings, stop at
 * |value| < 1e-8, TRIANGLE_MESH.cpp:2232-2271); vertex numbering, triangle
 * order and the degenerate filter are the linear mesher's. params: the
 * mesher uses escape 2000 and the black-hole stop (TRIANGLE_MESH.cpp:97).
 * precision selects the slab evaluation (QJ_FP32_BAND or QJ_FP64); vertices are
 * always FP64 (with QJ_FP32_BAND and iso 0 the end values are the slab's, whose
 * signs are the FP64 map's; the midpoints are always evaluated). Returns
 * QJ_ERR_NUMERIC (mesh still available) when an edge's end values do not
 * straddle the isosurface, where the reference exits. The next chunk's slab is
 * evaluated on a stream of the workspace's own while the current chunk is
 * meshed on `stream` (two slab buffers); that stream first waits for the work
 * queued on `stream` before the call, and `stream` waits for every slab.
 * Results are read with qj_mc_copy / qj_mc_device_ptrs. Synchronizes `stream`. */
qj_status qj_nlmc_run(qj_workspace* ws, const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                      double iso, qj_precision precision, double band, int32_t chunk_slices,
                      int64_t* n_vertices, int64_t* n_triangles, void* stream);

/* The non-linear mesher on a given field: TRIANGLE_MESH(const FIELD_3D& field,
 * top, bottom, expScaling, maxIterations, slice, isosurface)
 * (TRIANGLE_MESH.h:45, TRIANGLE_MESH.cpp:62-79 -> computeNonlinearMarchingCubes
 * :152-215, computeNonlinearEdgeInterpolations :1824-1910). Cases come from
 * d_field (device, x fastest) with the float test (float)v > 0 (no
 * isovalue); each vertex is placed by the bisection of qj_nlmc_run on the map
 * re-evaluated in FP64 at both edge ends; vertex numbering, triangle order
 * and the degenerate filter are the linear mesher's. params as for
 * qj_nlmc_run (the reference's escape is 2000; the black-hole stop always
 * applies). Results via qj_mc_copy; the edge cache via qj_nlmc_cache.
 * Synchronizes `stream`. Returns QJ_ERR_NUMERIC (mesh still available) when
 * an edge's re-evaluated end values do not straddle the isosurface. */
qj_status qj_nlmc_field(qj_workspace* ws, const double* d_field, const qj_grid* grid, const qj_rational* f,
                        const qj_map_params* params, double iso, int64_t* n_vertices, int64_t* n_triangles,
                        void* stream);

/* The edge cache of the last qj_nlmc_run / qj_nlmc_resume on ws, the data of
 * the mesher's cache file (writeEdgeCacheHuge / readEdgeCacheHuge,
 * TRIANGLE_MESH.cpp:431-586): one edge per vertex, in vertex order, as the
 * lower endpoint's (x, y, z) and the axis (0 x, 1 y, 2 z), and the active
 * cells in cell order as (x, y, z, flag). Arrays are int32 x 4 per entry, host
 * or device (hipMemcpyDefault); pass null arrays to read the counts only. */
qj_status qj_nlmc_cache(qj_workspace* ws, int64_t* n_edges, int32_t* edges, int64_t* n_cells, int32_t* cells,
                        void* stream);

/* The mesher's resume path (TRIANGLE_MESH.cpp:703-711 with a cache present):
 * rebuild the mesh from an edge cache (host arrays in the qj_nlmc_cache
 * layout) without evaluating slabs: each listed edge's vertex by the same
 * bisection as qj_nlmc_run, triangles from the listed cells. Edges must be in
 * vertex order and inside the grid; a triangle whose edge is not listed is an
 * error. The mesh is identical to the run that wrote the cache. Results via
 * qj_mc_copy. Synchronizes `stream`. */
qj_status qj_nlmc_resume(qj_workspace* ws, const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                         double iso, int64_t n_edges, const int32_t* edges, int64_t n_cells, const int32_t* cells,
                         int64_t* n_vertices, int64_t* n_triangles, void* stream);

/* ---- FIELD_3D clean-up before meshing (the reference runs
 *      clampInfsToMean then setBorders(totalCells()) on a fractal before
 *      TRIANGLE_MESH(field), toothOptimizer.cpp:250-254) ---- */

/* FIELD_3D::clampInfsToMean (FIELD_3D.cpp:5182-5207): every infinite cell of
 * d_field (device, n doubles) takes the mean of the non-infinite cells. NaN
 * cells are summed like the reference's serial sum does, so one NaN makes the
 * mean NaN; with no summed cell the mean is 0/0 = NaN. The finite cells are
 * summed in double-double arithmetic in a fixed order, so the mean is the
 * correctly rounded sum over the count (the reference's serial sum differs
 * from it by at most (n-1) 2^-53 sum|x|). Enqueued on `stream`; when mean_out
 * (host) is not NULL the call synchronizes and stores the mean there. */
qj_status qj_field_clamp_infs_to_mean(qj_workspace* ws, double* d_field, int64_t n, double* mean_out,
                                      void* stream);
/* FIELD_3D::clampNansToMean (FIELD_3D.cpp:5212-5234): the same for NaN cells,
 * with the mean of the non-NaN cells (an infinite cell makes it infinite). */
qj_status qj_field_clamp_nans_to_mean(qj_workspace* ws, double* d_field, int64_t n, double* mean_out,
                                      void* stream);
/* FIELD_3D::setBorders (FIELD_3D.cpp:5357-5382): the cells on the six faces of
 * the grid take `value`. Enqueued on `stream`. */
qj_status qj_field_set_borders(const qj_grid* grid, double value, double* d_field, void* stream);

/* FIELD_3D::fieldMin / fieldMax (FIELD_3D.cpp:2960-2971, 3021-3032) of d_field
 * (device, n >= 1 doubles), with the serial scan's behaviour: NaN cells are
 * skipped unless cell 0 is NaN (then both are NaN), and of equal values the
 * first cell's wins. Synchronizes `stream`; results on the host. */
qj_status qj_field_min_max(qj_workspace* ws, const double* d_field, int64_t n, double* min_out, double* max_out,
                           void* stream);

/* FIELD_3D::invertedDistance (FIELD_3D.cpp:5241-5297) of a signed distance
 * field d_dist on `grid` (device): the cells with a 26-neighbour of the other
 * sign (outerShell + innerShell, :5070-5145) get 1, every other cell 1 + its
 * 6-neighbour layer number (sweep after sweep, each reading the previous
 * one), then 1 / number, negated inside (distance < 0) and quartered outside.
 * normalized != 0: invertedDistanceNormalized (:5316-5352), the result over
 * the sum of its absolute values, then the inside and the outside each scaled
 * to absolute sum 0.5. abs_sum (host, may be NULL): the absolute sum of the
 * result (OPTIMIZE_3D::read's _inverseFieldSum, OPTIMIZE_3D.cpp:3375-3378).
 * Sums in double-double (the reference's serial sums differ by at most
 * (n - 1) 2^-53 of the absolute sum). A grid where some cell can never be
 * numbered (no sign change anywhere: the reference sweeps forever) returns
 * QJ_ERR_NUMERIC. Synchronizes `stream`. */
qj_status qj_field_inverted_distance(qj_workspace* ws, const qj_grid* grid, const double* d_dist, int32_t normalized,
                                     double* d_out, double* abs_sum, void* stream);

/* OPTIMIZE_3D::weightScore (OPTIMIZE_3D.cpp:4502-4517) and
 * inverseDistanceScore (:4522-4563) for n_thresholds thresholds t (host
 * array): scores[k] = sum over the cells of `grid` of
 * -tanh(tanh_alpha (F - t_k)) w(cell centre), with w the trilinear lookup of
 * the weight field (wgrid, d_weight; FIELD_3D.cpp:966-1029). max_score (host,
 * may be NULL): computeMaximumScaledPowerRationalScore (:4484-4498), the sum
 * of |w(cell centre)|. One launch for all thresholds (at most 65534); sums
 * in double-double. Synchronizes `stream`. */
qj_status qj_weight_scores(qj_workspace* ws, const qj_grid* grid, const double* d_field, const qj_grid* wgrid,
                           const double* d_weight, double tanh_alpha, const double* thresholds, int32_t n_thresholds,
                           double* scores, double* max_score, void* stream);

/* ---- gradients ---- */

/* One-iteration AP score and exponent gradients (max_iter must be 1, as the
 * optimizer forces at toothOptimizer.cpp:402):
 *   score        = sum_cells -tanh(alpha F) w                (weightScore, OPTIMIZE_3D.cpp:4502-4517)
 *   grad_exp     = sum -alpha w (1 - tanh^2(alpha (F - thr)))  (computeExpGradientAP, :3589-3625)
 *   grad_top[i]  = d score / d top power i        (computeTopGradientAP, :8905-9106)
 *   grad_bottom[j] = reference bottom formula     (computeBottomGradientAP, :10012)
 *   grad_top_scalar / grad_bottom_scalar           (:9250, :10360)
 * w(cell) is the trilinear lookup of the weight field (wgrid, d_weight) at the
 * cell centre (FIELD_3D.cpp:966-1029). d_field (device) receives the AP
 * field. Results are written to host memory; the call synchronizes `stream`. */
typedef struct qj_ap_grads {
  double score;
  double grad_exp;
  double grad_top_scalar;
  double grad_bottom_scalar;
  int64_t adaptive_cells;
} qj_ap_grads;
qj_status qj_ap_score_gradients(qj_workspace* ws, const qj_grid* grid, const qj_rational* f,
                                const qj_map_params* params, const qj_grid* wgrid, const double* d_weight,
                                double tanh_alpha, double tanh_threshold, double* d_field,
                                double* grad_top, double* grad_bottom, qj_ap_grads* out, void* stream);

/* Reverse-mode adjoint of the map at points, recomputing each iteration
 * instead of storing trajectories. Given dJ/dF per point (device, n doubles),
 * accumulates dJ/d(root components) into d_grad_roots ((n_top + n_bottom) x 4,
 * top first, component order w,x,y,z: the packing of
 * computeLogRationalCenteredGradient, OPTIMIZE_3D.cpp:2872-2951) and
 * dJ/d(powers) into d_grad_powers (n_top + n_bottom). Both are overwritten.
 * precision QJ_FP64 or QJ_FP32_BAND (FP32 adjoint; band ignored). */
qj_status qj_points_adjoint(qj_workspace* ws, const double* d_points, int64_t n, const qj_rational* f,
                            const qj_map_params* params, qj_precision precision, const double* d_dJdF,
                            double* d_grad_roots, double* d_grad_powers, void* stream);

/* ---- request loop (continuous batching of field solves) ----
 * Requests of any grid size and root count are admitted while their HBM
 * footprint fits the budget, split into fixed-size voxel chunks, and packed
 * with chunks of other requests into launches of a hipGraph captured once
 * per launch shape; results stay resident in device memory. */
typedef struct qj_server qj_server;

typedef struct qj_server_config {
  int64_t hbm_budget_bytes;   /* admission budget for fields + roots (0 = 8 GiB) */
  int32_t chunk_voxels;       /* voxels per chunk (0 = 65536) */
  int32_t chunks_per_launch;  /* chunks per graph launch (0 = 512) */
  int32_t max_roots;          /* largest root count per polynomial (0 = 256) */
  int32_t inflight_launches;  /* launches in flight (0 = 2) */
  int32_t precision;          /* qj_precision */
  double band;
  int32_t device;
  double launch_work;         /* root-term iterations per launch (voxels x max_iter x roots), so every
                                 launch takes about the same GPU time (0 = 3e9, about 1.5 ms) */
  double max_pending_work;    /* qj_server_submit refuses a request with QJ_ERR_BUSY when the root-term
                                 iterations queued ahead of it in shortest-first order (the outstanding
                                 requests with no more work than it, its own included) exceed this
                                 (0 = no limit; a request arriving at an idle server is always taken) */
  double aging_s;             /* a request waiting longer than this many seconds is admitted before the
                                 others (oldest first; while it does not fit, requests of a quarter of its
                                 footprint or more wait behind it) and takes up to a quarter of each
                                 launch (0 = 0.25) */
  int32_t check_samples;      /* per-result error statement: cells of every finished field re-evaluated with
                                 the reference method in FP64 and compared (qj_completion.checked ...; 0 = off) */
  double escalate_abs_dF;     /* FP32_BAND servers with check_samples: a result whose sample shows a sign flip,
                                 a class mismatch or max |F - F_fp64| above this is computed again with every
                                 voxel in FP64 (exact integer powers) before it completes; its completion says
                                 escalated = 1 and carries the second cross-check (0 = off) */
} qj_server_config;

typedef struct qj_request {
  qj_grid grid;
  qj_poly top;
  qj_poly bottom;
  qj_map_params params;
  int32_t keep_field;        /* 1: keep the field resident until qj_server_release */
  double arrival_time;       /* caller clock (seconds), reported back */
} qj_request;

typedef struct qj_completion {
  uint64_t id;
  qj_status status;
  double arrival_time;
  double admit_time;        /* server clock (qj_server_now) */
  double done_time;
  const double* d_field;    /* device field if keep_field, else NULL */
  int64_t voxels;
  /* the per-result cross-check (qj_server_config.check_samples; zeros when off),
     as in qj_eval_stats: cells compared, sign disagreements of (float)F > 0,
     class disagreements, largest absolute and relative |F - F_fp64| */
  int64_t checked;
  int64_t check_sign_flips;
  int64_t check_class_mismatch;
  double check_max_abs_dF;
  double check_max_rel_dF;
  int32_t escalated;        /* 1: recomputed in FP64 after its FP32 cross-check (escalate_abs_dF) */
} qj_completion;

qj_status qj_server_create(const qj_server_config* cfg, qj_server** out);
double qj_server_now(qj_server* s); /* seconds on the server's monotonic clock */
/* Copies the request (roots included); never blocks on the GPU. Returns
 * QJ_ERR_BUSY (no id, nothing queued) when the work queued ahead of it would
 * exceed max_pending_work. */
qj_status qj_server_submit(qj_server* s, const qj_request* req, uint64_t* id);
/* Up to max completions; waits up to timeout_s for at least one (0 = poll). */
qj_status qj_server_poll(qj_server* s, qj_completion* out, int32_t max, int32_t* n_out, double timeout_s);
/* Copies a kept field (completion->voxels doubles, x fastest) to dst (host or
 * device); synchronous. Call before qj_server_release of that request. */
qj_status qj_server_fetch(qj_server* s, const qj_completion* c, double* dst, void* stream);
qj_status qj_server_release(qj_server* s, uint64_t id); /* frees a kept field */
qj_status qj_server_destroy(qj_server* s);

#ifdef __cplusplus
}
#endif
#endif
