/*
 * Oracle: CPU restatement (FP64, plus x87 long double for the adaptive
 * re-pass) of QUIJIBO's rational-quaternion Julia field path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the shipped library links, loads or
 * calls this code. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, and only to compare against / as the CPU baseline.
 *
 * Every function cites the reference file:line it restates. Quaternion slot
 * order is the reference's QUATERNION(w, x, y, z) (QUATERNION.h:36,150-154).
 */
#ifndef QO_H
#define QO_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct qo_grid {
  int32_t res[3];
  double center[3];
  double lengths[3];
  double rotation[4]; /* quaternion (w,x,y,z); identity = (1,0,0,0) (FIELD_3D.cpp:30) */
} qo_grid;

typedef struct qo_poly {
  int32_t n;            /* _totalRoots */
  const double* roots;  /* n x 4 (w,x,y,z) */
  const double* powers; /* n, _rootPowers */
  double power_scalar;  /* _powerScalar */
} qo_poly;

enum {
  QO_BLACK_HOLE = 1u,  /* stop when |q| < 10*FLT_MIN (OPTIMIZE_3D.cpp:1358-1359) */
  QO_SKIP_ORIGIN = 2u, /* leave the cell unwritten when |q0| < 1e-8 (OPTIMIZE_3D.cpp:1428) */
  QO_AP_REPASS = 4u,   /* overflow flags + long double re-pass (OPTIMIZE_3D.cpp:1440-1567) */
  QO_LAGRANGE = 8u,    /* RATIONAL_4D::evaluateLagrange order (RATIONAL_4D.cpp:27-88) */
  QO_EXTENDED = 16u    /* the default build's Real = long double (SETTINGS.h:10) for the whole map */
};

typedef struct qo_params {
  int32_t max_iter;    /* _maxIterations */
  double escape;       /* 2000 (OPTIMIZE_3D.cpp:1291) or 20 (:1388,:1589) */
  double exp_scaling;  /* _expScaling */
  double slice;        /* _quaternionSlice, the k component of every voxel */
  uint32_t flags;
} qo_params;

/* quaternion primitives on double[4] (w,x,y,z); exposed for the algebra tests */
void qo_qmul(const double* l, const double* r, double* out);
void qo_qdiv(const double* l, const double* r, double* out);
void qo_qpow(const double* q, double e, double* out);
void qo_qlog(const double* q, double* out);
double qo_qmag(const double* q);
void qo_poly_eval(const qo_poly* p, const double* q, double* out);

void qo_cell_center(const qo_grid* g, int x, int y, int z, double* out3);

/* Field map over a grid, x fastest. iters (optional) receives executed
 * iterations per cell; overflowed (optional) receives the AP overflow flag. */
void qo_field_map(const qo_grid* g, const qo_poly* top, const qo_poly* bottom,
                  const qo_params* p, double* field, int32_t* iters,
                  uint8_t* overflowed, int nthreads);

void qo_field_map_zrange(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p,
                         int z0, int z1, double* field, int32_t* iters, uint8_t* overflowed, int nthreads);

/* Same iteration at arbitrary spatial points (X,Y,Z) (slice from params). */
void qo_points_eval(const double* pts, int64_t n, const qo_poly* top,
                    const qo_poly* bottom, const qo_params* p, double* out,
                    int nthreads);

/* Linear marching cubes (TRIANGLE_MESH.cpp:989-1114). Returns 0 on success.
 * Call with vertices/triangles NULL to get counts; then with buffers sized
 * to the counts. */
int qo_marching_cubes(const qo_grid* g, const double* field, int64_t* n_vertices,
                      int64_t* n_triangles, double* vertices, int32_t* triangles);

/* Non-linear mesher (TRIANGLE_MESH.cpp:681-768): cases from log|q|, vertices by
 * bisection on log|q| - iso. Returns the count of edges whose end values do not
 * straddle zero (the reference exits on the first). Two-call protocol as above. */
int64_t qo_nonlinear_mc(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p, double iso,
                        int64_t* n_vertices, int64_t* n_triangles, double* vertices, int32_t* triangles);

/* The same with the cases from a given field (TRIANGLE_MESH.cpp:62-79, 152-215);
 * given == NULL evaluates the map's own field. */
int64_t qo_nonlinear_mc_field(const qo_grid* g, const double* given, const qo_poly* top, const qo_poly* bottom,
                              const qo_params* p, double iso, int64_t* n_vertices, int64_t* n_triangles,
                              double* vertices, int32_t* triangles);

/* FIELD_3D::operator()(VEC3F) trilinear lookup (FIELD_3D.cpp:966-1029). */
double qo_field_lookup(const qo_grid* g, const double* data, const double* pos3);

/* FIELD_3D::clampInfsToMean / clampNansToMean (FIELD_3D.cpp:5182-5234):
 * in place, serial sum; return the mean used. */
double qo_clamp_infs_to_mean(double* data, int64_t n);
double qo_clamp_nans_to_mean(double* data, int64_t n);
/* FIELD_3D::setBorders (FIELD_3D.cpp:5357-5382) */
void qo_set_borders(const qo_grid* g, double* data, double value);

/* FIELD_3D::invertedDistance / invertedDistanceNormalized (FIELD_3D.cpp:5241-5352)
 * of a signed distance field; 0, or -1 when a cell can never be reached. */
int qo_inverted_distance(const qo_grid* g, const double* dist, double* out);
int qo_inverted_distance_normalized(const qo_grid* g, const double* dist, double* out, double* abs_sum);
/* weightScore / inverseDistanceScore (threshold) and the maximum score
 * (OPTIMIZE_3D.cpp:4484-4563) */
void qo_weight_score(const qo_grid* g, const double* field, const qo_grid* wg, const double* weight, double alpha,
                     double threshold, double* score, double* max_score);

/* Score and analytic gradients at one iteration (AP semantics).
 * weight grid/data: the weight field sampled at each fractal cell centre. */
typedef struct qo_ap_result {
  double score;           /* weightScore (OPTIMIZE_3D.cpp:4502-4517) */
  double grad_exp;        /* computeExpGradientAP (:3589-3625) */
  double grad_top_scalar; /* computeTopBulkScalarGradientAP (:9250) */
  double grad_bottom_scalar; /* computeBottomBulkScalarGradientAP (:10360) */
  int64_t overflowed;
} qo_ap_result;

void qo_ap_score_gradients(const qo_grid* g, const qo_poly* top, const qo_poly* bottom,
                           const qo_params* p, const qo_grid* wg, const double* weight,
                           double tanh_alpha, double tanh_threshold, double* field,
                           double* grad_top, double* grad_bottom, qo_ap_result* res,
                           int nthreads);

#ifdef __cplusplus
}
#endif
#endif
