// Host side of the C ABI: handles, argument checking, launch plumbing.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>

#include "qj_internal.h"

namespace qj {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

qj_status hip_fail(hipError_t e, const char* what) {
  set_error("%s failed: %s", what, hipGetErrorString(e));
  return e == hipErrorOutOfMemory ? QJ_ERR_NOMEM : QJ_ERR_HIP;
}

// QUATERNION::toRotationMatrix (QUATERNION.cpp:395-412) evaluated on the host
// in the reference's order; the host is built without FMA contraction.
void rotation_matrix(const double* q, double* m) {
  const double w = q[0], x = q[1], y = q[2], z = q[3];
  m[0] = 1 - 2 * y * y - 2 * z * z;
  m[1] = 2 * x * y - 2 * w * z;
  m[2] = 2 * x * z + 2 * w * y;
  m[3] = 2 * x * y + 2 * w * z;
  m[4] = 1 - 2 * x * x - 2 * z * z;
  m[5] = 2 * y * z - 2 * w * x;
  m[6] = 2 * x * z - 2 * w * y;
  m[7] = 2 * y * z + 2 * w * x;
  m[8] = 1 - 2 * x * x - 2 * y * y;
}

void make_geom(const qj_grid* g, GridGeom* o) {
  for (int i = 0; i < 3; i++) {
    o->res[i] = g->res[i];
    o->dx[i] = g->lengths[i] / g->res[i];
    o->lo[i] = g->center[i] - 0.5 * g->lengths[i];
    o->center[i] = g->center[i];
    o->lengths[i] = g->lengths[i];
  }
  rotation_matrix(g->rotation, o->rot);
  const uint64_t xy = (uint64_t)g->res[0] * (uint64_t)g->res[1];
  o->div_x = make_fastdiv((uint32_t)(g->res[0] > 0 ? g->res[0] : 1));
  o->div_xy = make_fastdiv(xy > 0 && xy <= 0xffffffffull ? (uint32_t)xy : 1u);
}

double near_bits(const std::vector<RootRec64>& roots) {
  double rmax = 0.0, emax = 0.0;
  for (const RootRec64& r : roots) {
    rmax = fmax(rmax, sqrt(r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z));
    emax = fmax(emax, fabs(r.e));
  }
  return emax * log2(2.0 * rmax + 1.0) + 1.0;
}

double term_bits(const std::vector<RootRec64>& roots, double escape) {
  double tb = 0.0;
  for (const RootRec64& r : roots) {
    const double rm = sqrt(r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z);
    const double b = fabs(r.e) * log2(fmax(escape, 1.0) + rm + 1.0);
    tb = fmax(tb, b);
  }
  return tb + 2.0;
}

qj_status make_map_args(const qj_grid* g, const qj_rational* f, const qj_map_params* p, MapArgs* a) {
  if (!g || !f || !p || !a) {
    set_error("null argument");
    return QJ_ERR_INVALID;
  }
  for (int i = 0; i < 3; i++) {
    if (g->res[i] < 1) {
      set_error("grid resolution must be >= 1 (got %d)", g->res[i]);
      return QJ_ERR_INVALID;
    }
    a->res[i] = g->res[i];
    // FIELD_3D ctor: _dx = _lengths[0] / _xRes (FIELD_3D.cpp:50-52)
    a->dx[i] = g->lengths[i] / g->res[i];
    // cellCenter: _center - (Real)0.5 * _lengths (FIELD_3D.cpp:820-823)
    a->lo[i] = g->center[i] - 0.5 * g->lengths[i];
    a->center[i] = g->center[i];
  }
  rotation_matrix(g->rotation, a->rot);
  if (p->max_iter < 0) {
    set_error("max_iter must be >= 0");
    return QJ_ERR_INVALID;
  }
  if ((p->flags & QJ_MAP_LAGRANGE) && (p->flags & (QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE))) {
    set_error("QJ_MAP_LAGRANGE does not combine with QJ_MAP_SKIP_ORIGIN or QJ_MAP_ADAPTIVE");
    return QJ_ERR_INVALID;
  }
  if ((p->flags & QJ_MAP_EXTENDED_RANGE) && (p->flags & (QJ_MAP_LAGRANGE | QJ_MAP_ADAPTIVE))) {
    set_error("QJ_MAP_EXTENDED_RANGE does not combine with QJ_MAP_LAGRANGE or QJ_MAP_ADAPTIVE");
    return QJ_ERR_INVALID;
  }
  a->max_iter = p->max_iter;
  a->zoff = 0;
  a->gen = 0;
  for (const RootRec64& r : f->h64) a->gen |= (r.e == 1.0 || r.e == 2.0) ? 0 : 1;
  a->slice = p->slice;
  a->exp_scaling = p->exp_scaling;
  a->escape = p->escape;
  a->flags = p->flags;
  a->nt = f->nt;
  a->nb = f->nb;
  a->sc_top = f->sc_top;
  a->sc_bottom = f->sc_bottom;
  a->term_bits = term_bits(f->h64, p->escape);
  a->near_bits = near_bits(f->h64);
  return QJ_OK;
}

// The lists of an evaluation of n cells: a band list of a quarter of the
// cells (at least 2^16, at most n; cells that need FP64 beyond it are marked
// in the output and swept) and, for QJ_MAP_ADAPTIVE, an adaptive list of n.
// At 512^3 that is 128 MiB of band list against the 1 GiB field.
qj_status ensure_list(qj_workspace* ws, int64_t n, bool adaptive, hipStream_t s) {
  const int64_t band = std::min(n, std::max(n / 4, (int64_t)1 << 16));
  const int64_t adapt = adaptive ? n : 0;
  if (band <= ws->list_cap && adapt <= ws->alist_cap && ws->d_list) return QJ_OK;
  const int64_t nb = std::max(band, ws->list_cap), na = std::max(adapt, ws->alist_cap);
  if (ws->d_list) {
    QJ_HIP(hipStreamSynchronize(s));
    QJ_HIP(hipFree(ws->d_list));
    ws->d_list = ws->d_alist = nullptr;
    ws->list_cap = ws->alist_cap = 0;
  }
  QJ_HIP(hipMalloc(&ws->d_list, sizeof(uint32_t) * (size_t)(nb + na + 1)));
  ws->d_alist = ws->d_list + nb;
  ws->list_cap = nb;
  ws->alist_cap = na;
  return QJ_OK;
}

qj_status ensure_partials(qj_workspace* ws, int64_t n, hipStream_t s) {
  if (n <= ws->partials_cap) return QJ_OK;
  if (ws->d_partials) {
    QJ_HIP(hipStreamSynchronize(s));
    QJ_HIP(hipFree(ws->d_partials));
    ws->d_partials = nullptr;
    ws->partials_cap = 0;
  }
  QJ_HIP(hipMalloc(&ws->d_partials, sizeof(double) * (size_t)n));
  ws->partials_cap = n;
  return QJ_OK;
}

void timing_mark(qj_workspace* ws, hipStream_t s) {
  if (!ws->timing) return;
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  hipEventRecord(e, s);
  ws->ev.push_back(e);
}

static qj_status check_poly(const qj_poly* p, const char* which, bool allow_empty) {
  if (!p) {
    set_error("%s polynomial is null", which);
    return QJ_ERR_INVALID;
  }
  if (p->n_roots < 0 || (!allow_empty && p->n_roots < 1)) {
    set_error("%s polynomial needs %s roots (got %d)", which, allow_empty ? ">= 0" : ">= 1", p->n_roots);
    return QJ_ERR_INVALID;
  }
  if (p->n_roots > 0 && (!p->roots || !p->powers)) {
    set_error("%s polynomial has null roots/powers", which);
    return QJ_ERR_INVALID;
  }
  return QJ_OK;
}

static void fill_records(const qj_poly* p, std::vector<RootRec64>& h64, std::vector<RootRec32>& h32) {
  for (int i = 0; i < p->n_roots; i++) {
    RootRec64 r;
    r.w = p->roots[4 * i + 0];
    r.x = p->roots[4 * i + 1];
    r.y = p->roots[4 * i + 2];
    r.z = p->roots[4 * i + 3];
    r.p = p->powers[i];
    r.e = p->power_scalar * p->powers[i];  // _powerScalar * _rootPowers[x], POLYNOMIAL_4D.cpp:914,919
    h64.push_back(r);
    RootRec32 t;
    t.w = (float)r.w;
    t.x = (float)r.x;
    t.y = (float)r.y;
    t.z = (float)r.z;
    t.e = (float)r.e;
    t.kind = (r.e == 1.0) ? 1 : (r.e == 2.0 ? 2 : 0);
    h32.push_back(t);
  }
}

void pack_roots32(const RootRec32* rec, int nt, int nb, std::vector<uint8_t>& blk) {
  const int n = nt + nb;
  blk.assign(roots32_bytes(n), 0);
  F4* q = (F4*)blk.data();
  int32_t* pat = (int32_t*)(q + n);
  float* e = (float*)(pat + n);
  for (int i = 0; i < n; i++) {
    q[i] = F4{rec[i].w, rec[i].x, rec[i].y, rec[i].z};
    e[i] = rec[i].e;
    // kinds of roots i..i+3 of the same polynomial
    const int end = i < nt ? nt : n;
    int32_t p = 0;
    for (int j = 0; j < 4 && i + j < end; j++) {
      const int k = rec[i + j].kind;
      p |= ((k == 2 ? kPatSquare : 0) | (k == 0 ? kPatGeneral : 0)) << (2 * j);
    }
    pat[i] = p;
  }
}

}  // namespace qj

using namespace qj;

extern "C" {

const char* qj_version(void) { return "quijibo-mi355x 0.1 (gfx950)"; }
const char* qj_last_error(void) { return g_err; }

const char* qj_status_string(qj_status s) {
  switch (s) {
    case QJ_OK: return "ok";
    case QJ_ERR_INVALID: return "invalid argument";
    case QJ_ERR_HIP: return "HIP runtime error";
    case QJ_ERR_NOMEM: return "out of device memory";
    case QJ_ERR_CAPACITY: return "capacity exceeded";
    case QJ_ERR_NUMERIC: return "inconsistent numeric state";
    case QJ_ERR_SHUTDOWN: return "server shutting down";
    case QJ_ERR_BUSY: return "refused: pending work over the limit";
    default: return "unknown status";
  }
}

qj_status qj_init(int32_t device) {
  int n = 0;
  QJ_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) {
    set_error("device %d not present (%d devices)", device, n);
    return QJ_ERR_INVALID;
  }
  QJ_HIP(hipSetDevice(device));
  // the allocation-free entry point's scratch address on this GPU
  return scratch_setup(device, nullptr);
}

qj_status qj_rational_create(const qj_poly* top, const qj_poly* bottom, qj_rational** out) {
  if (!out) return QJ_ERR_INVALID;
  *out = nullptr;
  qj_status st = check_poly(top, "top", false);
  if (st) return st;
  qj_poly empty{0, nullptr, nullptr, 1.0};
  if (!bottom) bottom = &empty;
  st = check_poly(bottom, "bottom", true);
  if (st) return st;
  qj_rational* f = new (std::nothrow) qj_rational();
  if (!f) return QJ_ERR_NOMEM;
  f->nt = top->n_roots;
  f->nb = bottom->n_roots;
  f->sc_top = top->power_scalar;
  f->sc_bottom = bottom->power_scalar;
  fill_records(top, f->h64, f->h32);
  fill_records(bottom, f->h64, f->h32);
  pack_roots32(f->h32.data(), f->nt, f->nb, f->blk32);
  hipGetDevice(&f->device);
  const size_t n = f->h64.size();
  hipError_t e = hipMalloc(&f->d64, sizeof(RootRec64) * n);
  if (e == hipSuccess) e = hipMalloc(&f->d32, f->blk32.size());
  if (e == hipSuccess) e = hipMemcpy(f->d64, f->h64.data(), sizeof(RootRec64) * n, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(f->d32, f->blk32.data(), f->blk32.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    hipFree(f->d64);
    hipFree(f->d32);
    delete f;
    return hip_fail(e, "qj_rational_create");
  }
  *out = f;
  return QJ_OK;
}

qj_status qj_rational_update(qj_rational* f, const qj_poly* top, const qj_poly* bottom, void* stream) {
  if (!f) return QJ_ERR_INVALID;
  qj_poly empty{0, nullptr, nullptr, 1.0};
  if (!bottom) bottom = &empty;
  qj_status st = check_poly(top, "top", false);
  if (st) return st;
  st = check_poly(bottom, "bottom", true);
  if (st) return st;
  if (top->n_roots != f->nt || bottom->n_roots != f->nb) {
    set_error("qj_rational_update: root counts changed (%d,%d) -> (%d,%d)", f->nt, f->nb, top->n_roots,
              bottom->n_roots);
    return QJ_ERR_INVALID;
  }
  hipStream_t s = (hipStream_t)stream;
  // the previous upload may still be reading the host staging copy
  QJ_HIP(hipStreamSynchronize(s));
  f->h64.clear();
  f->h32.clear();
  f->sc_top = top->power_scalar;
  f->sc_bottom = bottom->power_scalar;
  fill_records(top, f->h64, f->h32);
  fill_records(bottom, f->h64, f->h32);
  pack_roots32(f->h32.data(), f->nt, f->nb, f->blk32);
  const size_t n = f->h64.size();
  QJ_HIP(hipMemcpyAsync(f->d64, f->h64.data(), sizeof(RootRec64) * n, hipMemcpyHostToDevice, s));
  QJ_HIP(hipMemcpyAsync(f->d32, f->blk32.data(), f->blk32.size(), hipMemcpyHostToDevice, s));
  return QJ_OK;
}

qj_status qj_rational_counts(const qj_rational* f, int32_t* n_top, int32_t* n_bottom) {
  if (!f) return QJ_ERR_INVALID;
  if (n_top) *n_top = f->nt;
  if (n_bottom) *n_bottom = f->nb;
  return QJ_OK;
}

qj_status qj_rational_destroy(qj_rational* f) {
  if (!f) return QJ_OK;
  hipFree(f->d64);
  hipFree(f->d32);
  delete f;
  return QJ_OK;
}

qj_status qj_workspace_create(qj_workspace** out) {
  if (!out) return QJ_ERR_INVALID;
  qj_workspace* ws = new (std::nothrow) qj_workspace();
  if (!ws) return QJ_ERR_NOMEM;
  hipGetDevice(&ws->device);
  hipError_t e = hipMalloc(&ws->d_counters, sizeof(unsigned long long) * CNT_N * kCounterSpread);
  if (e == hipSuccess) e = hipMemset(ws->d_counters, 0, sizeof(unsigned long long) * CNT_N * kCounterSpread);
  if (e == hipSuccess) e = hipMalloc(&ws->d_list_count, 4 * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMemset(ws->d_list_count, 0, 4 * sizeof(uint32_t));
  if (e != hipSuccess) {
    hipFree(ws->d_counters);
    hipFree(ws->d_list_count);
    delete ws;
    return hip_fail(e, "qj_workspace_create");
  }
  *out = ws;
  return QJ_OK;
}

qj_status qj_workspace_destroy(qj_workspace* ws) {
  if (!ws) return QJ_OK;
  hipDeviceSynchronize();
  for (hipEvent_t e : ws->ev) hipEventDestroy(e);
  mc_free(ws);
  hipFree(ws->d_counters);
  hipFree(ws->d_list);
  hipFree(ws->d_list_count);
  hipFree(ws->d_partials);
  delete ws;
  return QJ_OK;
}

qj_status qj_workspace_stats(qj_workspace* ws, void* stream, qj_eval_stats* out) {
  if (!ws || !out) return QJ_ERR_INVALID;
  unsigned long long h[CNT_N * kCounterSpread];
  QJ_HIP(hipStreamSynchronize((hipStream_t)stream));
  QJ_HIP(hipMemcpy(h, ws->d_counters, sizeof(h), hipMemcpyDeviceToHost));
  int64_t t[CNT_N] = {0};
  for (int c = 0; c < CNT_N; c++)
    for (int i = 0; i < kCounterSpread; i++) t[c] += (int64_t)h[c * kCounterSpread + i];
  out->voxels = t[CNT_VOXELS];
  out->iterations = t[CNT_ITERS];
  out->band_voxels = t[CNT_BAND_VOXELS];
  out->band_iterations = t[CNT_BAND_ITERS];
  out->adaptive_voxels = t[CNT_ADAPTIVE];
  out->nonfinite = t[CNT_NONFINITE];
  out->checked = t[CNT_CHECKED];
  out->check_sign_flips = t[CNT_CHECK_FLIPS];
  out->check_class_mismatch = t[CNT_CHECK_CLASS];
  unsigned long long mabs = 0, mrel = 0;
  for (int i = 0; i < kCounterSpread; i++) {
    mabs = h[CNT_CHECK_MAXABS * kCounterSpread + i] > mabs ? h[CNT_CHECK_MAXABS * kCounterSpread + i] : mabs;
    mrel = h[CNT_CHECK_MAXREL * kCounterSpread + i] > mrel ? h[CNT_CHECK_MAXREL * kCounterSpread + i] : mrel;
  }
  memcpy(&out->check_max_abs_dF, &mabs, sizeof(double));
  memcpy(&out->check_max_rel_dF, &mrel, sizeof(double));
  return QJ_OK;
}

qj_status qj_workspace_check(qj_workspace* ws, int64_t samples) {
  if (!ws || samples < 0 || samples > (int64_t(1) << 24)) {
    set_error("qj_workspace_check: null workspace or samples outside [0, 2^24]");
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  ws->check_samples = samples;
  return QJ_OK;
}

qj_status qj_workspace_reset_stats(qj_workspace* ws, void* stream) {
  if (!ws) return QJ_ERR_INVALID;
  QJ_HIP(hipMemsetAsync(ws->d_counters, 0, sizeof(unsigned long long) * CNT_N * kCounterSpread,
                        (hipStream_t)stream));
  return QJ_OK;
}

qj_status qj_workspace_timing(qj_workspace* ws, int32_t enable) {
  if (!ws) return QJ_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ws->mu);
  for (hipEvent_t e : ws->ev) hipEventDestroy(e);
  ws->ev.clear();
  ws->timing = enable != 0;
  return QJ_OK;
}

qj_status qj_workspace_kernel_ms(qj_workspace* ws, double* bulk_ms, double* band_ms, int64_t* launches) {
  if (!ws) return QJ_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ws->mu);
  double b = 0, d = 0;
  const size_t n = ws->ev.size() / 4;
  for (size_t i = 0; i < n; i++) {
    float ms = 0;
    QJ_HIP(hipEventSynchronize(ws->ev[4 * i + 3]));
    QJ_HIP(hipEventElapsedTime(&ms, ws->ev[4 * i + 0], ws->ev[4 * i + 1]));
    b += ms;
    QJ_HIP(hipEventElapsedTime(&ms, ws->ev[4 * i + 2], ws->ev[4 * i + 3]));
    d += ms;
  }
  if (bulk_ms) *bulk_ms = b;
  if (band_ms) *band_ms = d;
  if (launches) *launches = (int64_t)n;
  return QJ_OK;
}

qj_status qj_field_eval(qj_workspace* ws, const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                        qj_precision precision, double band, double* d_field, void* stream) {
  if (!ws || !d_field) {
    set_error("qj_field_eval: null workspace or output");
    return QJ_ERR_INVALID;
  }
  if (precision != QJ_FP64 && precision != QJ_FP32_BAND && precision != QJ_FP64_EXACT) {
    set_error("qj_field_eval: unknown precision %d", (int)precision);
    return QJ_ERR_INVALID;
  }
  if (params && (params->flags & QJ_MAP_LAGRANGE) && precision != QJ_FP64) {
    set_error("qj_field_eval: QJ_MAP_LAGRANGE needs QJ_FP64");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  std::lock_guard<std::mutex> lock(ws->mu);
  return launch_field(ws, a, f, precision, band, d_field, (hipStream_t)stream);
}

qj_status qj_points_eval(qj_workspace* ws, const double* d_points, int64_t n, const qj_rational* f,
                         const qj_map_params* params, qj_precision precision, double band, double* d_out,
                         void* stream) {
  if (!ws || (n > 0 && (!d_points || !d_out)) || n < 0) {
    set_error("qj_points_eval: bad arguments");
    return QJ_ERR_INVALID;
  }
  if (precision != QJ_FP64 && precision != QJ_FP32_BAND && precision != QJ_FP64_EXACT) return QJ_ERR_INVALID;
  if (params && (params->flags & QJ_MAP_LAGRANGE) && precision != QJ_FP64) {
    set_error("qj_points_eval: QJ_MAP_LAGRANGE needs QJ_FP64");
    return QJ_ERR_INVALID;
  }
  qj_grid g{{1, 1, 1}, {0, 0, 0}, {1, 1, 1}, {1, 0, 0, 0}};
  MapArgs a;
  qj_status st = make_map_args(&g, f, params, &a);
  if (st) return st;
  std::lock_guard<std::mutex> lock(ws->mu);
  return launch_points(ws, a, f, precision, band, d_points, n, d_out, (hipStream_t)stream);
}

qj_status qj_mc_run(qj_workspace* ws, const double* d_field, const qj_grid* grid, double iso, int64_t* n_vertices,
                    int64_t* n_triangles, void* stream) {
  if (!ws || !d_field || !grid || !n_vertices || !n_triangles) {
    set_error("qj_mc_run: null argument");
    return QJ_ERR_INVALID;
  }
  for (int i = 0; i < 3; i++)
    if (grid->res[i] < 1) return QJ_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ws->mu);
  return mc_run(ws, d_field, grid, iso, n_vertices, n_triangles, (hipStream_t)stream);
}

qj_status qj_mc_run_tiles(qj_workspace* ws, int32_t n, const double* const* d_fields, const qj_grid* grids,
                          double iso, int64_t* n_vertices, int64_t* n_triangles, void* stream) {
  if (!ws || n < 0 || (n > 0 && (!d_fields || !grids || !n_vertices || !n_triangles))) {
    set_error("qj_mc_run_tiles: null argument or negative count");
    return QJ_ERR_INVALID;
  }
  for (int k = 0; k < n; k++)
    if (!d_fields[k]) {
      set_error("qj_mc_run_tiles: field %d is null", k);
      return QJ_ERR_INVALID;
    }
  std::lock_guard<std::mutex> lock(ws->mu);
  return mc_run_tiles(ws, n, d_fields, grids, iso, n_vertices, n_triangles, (hipStream_t)stream);
}

qj_status qj_mc_run_tiles_into(qj_workspace* ws, int32_t n, const double* const* d_fields, const qj_grid* grids,
                               double iso, double* d_vertices, int64_t vertex_capacity, int32_t* d_triangles,
                               int64_t triangle_capacity, int64_t* n_vertices, int64_t* n_triangles,
                               void* stream) {
  if (!ws || n < 0 || (n > 0 && (!d_fields || !grids || !n_vertices || !n_triangles)) || !d_vertices ||
      !d_triangles || vertex_capacity < 0 || triangle_capacity < 0) {
    set_error("qj_mc_run_tiles_into: null argument or negative count or capacity");
    return QJ_ERR_INVALID;
  }
  for (int k = 0; k < n; k++)
    if (!d_fields[k]) {
      set_error("qj_mc_run_tiles_into: field %d is null", k);
      return QJ_ERR_INVALID;
    }
  std::lock_guard<std::mutex> lock(ws->mu);
  return mc_run_tiles(ws, n, d_fields, grids, iso, n_vertices, n_triangles, (hipStream_t)stream, d_vertices,
                      vertex_capacity, d_triangles, triangle_capacity);
}

qj_status qj_nlmc_run(qj_workspace* ws, const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                      double iso, qj_precision precision, double band, int32_t chunk_slices, int64_t* n_vertices,
                      int64_t* n_triangles, void* stream) {
  if (!ws || !grid || !f || !params || !n_vertices || !n_triangles) {
    set_error("qj_nlmc_run: null argument");
    return QJ_ERR_INVALID;
  }
  if (precision != QJ_FP64 && precision != QJ_FP32_BAND) return QJ_ERR_INVALID;
  if (params->flags & (QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE)) {
    set_error("qj_nlmc_run: the mesher's map has no origin skip or adaptive re-pass");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  // the mesher's slabs and vertex values always stop at the black hole
  // (TRIANGLE_MESH.cpp:374-376, 1789-1791)
  a.flags |= QJ_MAP_BLACK_HOLE;
  std::lock_guard<std::mutex> lock(ws->mu);
  int64_t bad = 0;
  return nlmc_run(ws, a, f, iso, precision, band, chunk_slices, n_vertices, n_triangles, &bad,
                  (hipStream_t)stream);
}

qj_status qj_nlmc_field(qj_workspace* ws, const double* d_field, const qj_grid* grid, const qj_rational* f,
                        const qj_map_params* params, double iso, int64_t* n_vertices, int64_t* n_triangles,
                        void* stream) {
  if (!ws || !d_field || !grid || !f || !params || !n_vertices || !n_triangles) {
    set_error("qj_nlmc_field: null argument");
    return QJ_ERR_INVALID;
  }
  if (params->flags & (QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE | QJ_MAP_LAGRANGE)) {
    set_error("qj_nlmc_field: the mesher's map has no origin skip, adaptive re-pass or Lagrange order");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  a.flags |= QJ_MAP_BLACK_HOLE;  // nonlinearValue's black-hole stop (TRIANGLE_MESH.cpp:1789-1791)
  std::lock_guard<std::mutex> lock(ws->mu);
  int64_t bad = 0;
  return nlmc_run(ws, a, f, iso, QJ_FP64, 0.0, 64, n_vertices, n_triangles, &bad, (hipStream_t)stream, d_field);
}

qj_status qj_mc_copy(qj_workspace* ws, double* vertices, int32_t* triangles, void* stream) {
  if (!ws) return QJ_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ws->mu);
  return mc_copy(ws, vertices, triangles, (hipStream_t)stream);
}

qj_status qj_mc_device_ptrs(qj_workspace* ws, const double** d_vertices, const int32_t** d_triangles) {
  if (!ws) return QJ_ERR_INVALID;
  return mc_ptrs(ws, d_vertices, d_triangles);
}

qj_status qj_nlmc_cache(qj_workspace* ws, int64_t* n_edges, int32_t* edges, int64_t* n_cells, int32_t* cells,
                        void* stream) {
  if (!ws) {
    set_error("qj_nlmc_cache: null workspace");
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  return nlmc_cache(ws, n_edges, edges, n_cells, cells, (hipStream_t)stream);
}

qj_status qj_nlmc_resume(qj_workspace* ws, const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                         double iso, int64_t n_edges, const int32_t* edges, int64_t n_cells, const int32_t* cells,
                         int64_t* n_vertices, int64_t* n_triangles, void* stream) {
  if (!ws || !grid || !f || !params || !n_vertices || !n_triangles) {
    set_error("qj_nlmc_resume: null argument");
    return QJ_ERR_INVALID;
  }
  if (params->flags & (QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE)) {
    set_error("qj_nlmc_resume: the mesher's map has no origin skip or adaptive re-pass");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  // the mesher's value always stops at the black hole (TRIANGLE_MESH.cpp:1789-1791)
  a.flags |= QJ_MAP_BLACK_HOLE;
  std::lock_guard<std::mutex> lock(ws->mu);
  return nlmc_resume(ws, a, f, iso, n_edges, edges, n_cells, cells, n_vertices, n_triangles, (hipStream_t)stream);
}

static qj_status clamp_entry(qj_workspace* ws, double* d_field, int64_t n, bool nans, double* mean_out,
                             void* stream, const char* name) {
  if (!ws || (n > 0 && !d_field) || n < 0) {
    set_error("%s: null workspace or field, or negative size", name);
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  return field_clamp_to_mean(ws, d_field, n, nans, mean_out, (hipStream_t)stream);
}

qj_status qj_field_clamp_infs_to_mean(qj_workspace* ws, double* d_field, int64_t n, double* mean_out,
                                      void* stream) {
  return clamp_entry(ws, d_field, n, false, mean_out, stream, "qj_field_clamp_infs_to_mean");
}

qj_status qj_field_clamp_nans_to_mean(qj_workspace* ws, double* d_field, int64_t n, double* mean_out,
                                      void* stream) {
  return clamp_entry(ws, d_field, n, true, mean_out, stream, "qj_field_clamp_nans_to_mean");
}

qj_status qj_field_set_borders(const qj_grid* grid, double value, double* d_field, void* stream) {
  if (!grid || !d_field) {
    set_error("qj_field_set_borders: null argument");
    return QJ_ERR_INVALID;
  }
  for (int i = 0; i < 3; i++)
    if (grid->res[i] < 1) {
      set_error("qj_field_set_borders: grid resolution must be >= 1 (got %d)", grid->res[i]);
      return QJ_ERR_INVALID;
    }
  return field_set_borders(grid->res, value, d_field, (hipStream_t)stream);
}

static bool grid_ok(const qj_grid* g) {
  return g && g->res[0] >= 1 && g->res[1] >= 1 && g->res[2] >= 1;
}

qj_status qj_field_min_max(qj_workspace* ws, const double* d_field, int64_t n, double* min_out, double* max_out,
                           void* stream) {
  if (!ws || !d_field || !min_out || !max_out || n < 1) {
    set_error("qj_field_min_max: null argument or empty field");
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  return field_min_max(ws, d_field, n, min_out, max_out, (hipStream_t)stream);
}

qj_status qj_field_inverted_distance(qj_workspace* ws, const qj_grid* grid, const double* d_dist, int32_t normalized,
                                     double* d_out, double* abs_sum, void* stream) {
  if (!ws || !grid_ok(grid) || !d_dist || !d_out || d_dist == d_out) {
    set_error("qj_field_inverted_distance: null argument, empty grid, or output aliasing the distance field");
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  return field_inverted_distance(ws, d_dist, grid, normalized != 0, d_out, abs_sum, (hipStream_t)stream);
}

qj_status qj_weight_scores(qj_workspace* ws, const qj_grid* grid, const double* d_field, const qj_grid* wgrid,
                           const double* d_weight, double tanh_alpha, const double* thresholds, int32_t n_thresholds,
                           double* scores, double* max_score, void* stream) {
  if (!ws || !grid_ok(grid) || !grid_ok(wgrid) || !d_field || !d_weight || n_thresholds < 0 ||
      (n_thresholds > 0 && (!thresholds || !scores)) || n_thresholds > 65534) {
    set_error("qj_weight_scores: null argument, empty grid, or threshold count outside [0, 65534]");
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(ws->mu);
  return weight_scores(ws, d_field, grid, wgrid, d_weight, tanh_alpha, thresholds, n_thresholds, scores, max_score,
                       (hipStream_t)stream);
}

qj_status qj_ap_score_gradients(qj_workspace* ws, const qj_grid* grid, const qj_rational* f,
                                const qj_map_params* params, const qj_grid* wgrid, const double* d_weight,
                                double tanh_alpha, double tanh_threshold, double* d_field, double* grad_top,
                                double* grad_bottom, qj_ap_grads* out, void* stream) {
  if (!ws || !wgrid || !d_weight || !d_field || !grad_top || !out || (f && f->nb > 0 && !grad_bottom)) {
    set_error("qj_ap_score_gradients: null argument");
    return QJ_ERR_INVALID;
  }
  if (!params || params->max_iter != 1) {
    set_error("qj_ap_score_gradients: the AP gradients are defined for max_iter == 1 (toothOptimizer.cpp:402)");
    return QJ_ERR_INVALID;
  }
  if (params->flags & QJ_MAP_EXTENDED_RANGE) {
    set_error("qj_ap_score_gradients: the AP pass has its own extended-range re-pass; QJ_MAP_EXTENDED_RANGE is "
              "not supported");
    return QJ_ERR_INVALID;
  }
  if (params->flags & QJ_MAP_LAGRANGE) {
    set_error("qj_ap_score_gradients: the AP gradients are of the factored order; QJ_MAP_LAGRANGE is not supported");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  for (int i = 0; i < 3; i++)
    if (wgrid->res[i] < 1) return QJ_ERR_INVALID;
  std::lock_guard<std::mutex> lock(ws->mu);
  return ap_grads(ws, a, f, wgrid, d_weight, tanh_alpha, tanh_threshold, d_field, grad_top, grad_bottom, out,
                  (hipStream_t)stream);
}

qj_status qj_points_adjoint(qj_workspace* ws, const double* d_points, int64_t n, const qj_rational* f,
                            const qj_map_params* params, qj_precision precision, const double* d_dJdF,
                            double* d_grad_roots, double* d_grad_powers, void* stream) {
  if (!ws || n < 0 || (n > 0 && (!d_points || !d_dJdF)) || !d_grad_roots || !d_grad_powers) {
    set_error("qj_points_adjoint: bad arguments");
    return QJ_ERR_INVALID;
  }
  if (precision != QJ_FP64 && precision != QJ_FP32_BAND) return QJ_ERR_INVALID;
  if (params && (params->flags & QJ_MAP_EXTENDED_RANGE)) {
    set_error("qj_points_adjoint: the adjoint runs in FP64 / FP32 range; QJ_MAP_EXTENDED_RANGE is not supported");
    return QJ_ERR_INVALID;
  }
  if (params && (params->flags & QJ_MAP_LAGRANGE)) {
    set_error("qj_points_adjoint: the adjoint is of the factored order; QJ_MAP_LAGRANGE is not supported");
    return QJ_ERR_INVALID;
  }
  qj_grid g{{1, 1, 1}, {0, 0, 0}, {1, 1, 1}, {1, 0, 0, 0}};
  MapArgs a;
  qj_status st = make_map_args(&g, f, params, &a);
  if (st) return st;
  std::lock_guard<std::mutex> lock(ws->mu);
  return points_adjoint(ws, a, f, precision, d_points, n, d_dJdF, d_grad_roots, d_grad_powers,
                        (hipStream_t)stream);
}

}  // extern "C"
