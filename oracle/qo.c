/*
 * Oracle: CPU restatement of QUIJIBO's rational-quaternion Julia field path
 * (FP64; x87 long double only where the reference itself switches to it).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, to compare against and as the CPU baseline. The
 * shipped library never links or calls it.
 *
 * Built with -ffp-contract=off so every expression rounds in the reference's
 * source order (g++ on x86-64 without -mfma does the same).
 *
 * Parity pinning: the algebra is pinned by closed forms and group identities,
 * the map by an independent complex-plane and 2x2-complex-matrix restatement,
 * the marching-cubes table by the reference's own case table, the gradients by
 * central differences (tests/test_oracle_*.py). The transcendental functions
 * come from the host libm; results at the ulp level depend on it.
 */
#include "qo.h"

#include <fenv.h>
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "qo_mc_table.h"

typedef struct { double w, x, y, z; } Q;
typedef struct { long double w, x, y, z; } QE;

/* ---------------- quaternion algebra (double) ---------------- */

/* operator*(QUATERNION,QUATERNION), QUATERNION.cpp:154-162; operator*= is the
 * same expression with left = *this (QUATERNION.h:51-58). */
static inline Q qmul(Q l, Q r) {
  Q f;
  f.x = l.y * r.z - l.z * r.y + r.w * l.x + l.w * r.x;
  f.y = l.z * r.x - l.x * r.z + r.w * l.y + l.w * r.y;
  f.z = l.x * r.y - l.y * r.x + r.w * l.z + l.w * r.z;
  f.w = l.w * r.w - l.x * r.x - r.y * l.y - l.z * r.z;
  return f;
}

/* operator*(QUATERNION, Real), QUATERNION.h:183-187 */
static inline Q qscale(Q q, double s) {
  Q f = {q.w * s, q.x * s, q.y * s, q.z * s};
  return f;
}

static inline Q qsub(Q a, Q b) {
  Q f = {a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z};
  return f;
}

/* conjugate() / magnitudeSq via operator/(q, Real) = q * (1.0/r):
 * QUATERNION.cpp:167-172 (inverse), :174-182 (division), :196-200 */
static inline Q qinv(Q r) {
  double m = r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z;
  double inv = 1.0 / m;
  Q c = {r.w, -r.x, -r.y, -r.z};
  return qscale(c, inv);
}

/* operator/(QUATERNION,QUATERNION), QUATERNION.cpp:174-182: left * inverse */
static inline Q qdiv(Q l, Q r) { return qmul(l, qinv(r)); }

/* QUATERNION::magnitude, QUATERNION.h:93 */
static inline double qmag(Q q) { return sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z); }

/* QUATERNION::dot, QUATERNION.h:137 */
static inline double qdot(Q a, Q b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }

/* QUATERNION::pow, QUATERNION.cpp:375-393 */
static inline Q qpow(Q q, double e) {
  const double partial = q.x * q.x + q.y * q.y + q.z * q.z;
  const double qMagnitude = sqrt(partial + q.w * q.w);
  const double vMagnitude = sqrt(partial);
  const double vMagnitudeInv = (vMagnitude > 0.0) ? 1.0 / vMagnitude : 0.0;
  const double scale = e * acos(q.w / qMagnitude) * vMagnitudeInv;
  const double magnitude = scale * vMagnitude;
  const double magnitudeInv = (magnitude > 0.0) ? 1.0 / magnitude : 0.0;
  const double exps = exp(e * log(qMagnitude));
  const double scale2 = scale * exps * magnitudeInv * sin(magnitude);
  Q f = {exps * cos(magnitude), scale2 * q.x, scale2 * q.y, scale2 * q.z};
  return f;
}

/* TVEC3<double>::magnitude uses sqrtf (VEC3.h:101-103) */
static inline double vec3_magnitude(double a, double b, double c) {
  return (double)sqrtf((float)(a * a + b * b + c * c));
}

/* QUATERNION::log, QUATERNION.cpp:352-367 */
static inline Q qlog(Q q) {
  const double s = q.w;
  const double qMagnitude = qmag(q);
  const double vMagnitude = vec3_magnitude(q.x, q.y, q.z);
  double nx = 0, ny = 0, nz = 0;
  if (vMagnitude > 0) { nx = q.x / vMagnitude; ny = q.y / vMagnitude; nz = q.z / vMagnitude; }
  const double a = acos(s / qMagnitude);
  Q f = {log(qMagnitude), nx * a, ny * a, nz * a};
  return f;
}

/* ---------------- quaternion algebra (long double, QUATERNION_T<long double>) ---------------- */

static inline QE emul(QE l, QE r) {
  QE f;
  f.x = l.y * r.z - l.z * r.y + r.w * l.x + l.w * r.x;
  f.y = l.z * r.x - l.x * r.z + r.w * l.y + l.w * r.y;
  f.z = l.x * r.y - l.y * r.x + r.w * l.z + l.w * r.z;
  f.w = l.w * r.w - l.x * r.x - r.y * l.y - l.z * r.z;
  return f;
}
static inline QE escale(QE q, long double s) {
  QE f = {q.w * s, q.x * s, q.y * s, q.z * s};
  return f;
}
static inline QE esub(QE a, QE b) {
  QE f = {a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z};
  return f;
}
static inline QE einv(QE r) {
  long double m = r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z;
  long double inv = 1.0L / m;
  QE c = {r.w, -r.x, -r.y, -r.z};
  return escale(c, inv);
}
static inline QE ediv(QE l, QE r) { return emul(l, einv(r)); }
static inline long double emag(QE q) { return sqrtl(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z); }
static inline long double edot(QE a, QE b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
/* QUATERNION_T<Float>::pow, QUATERNION_T.h:466-484 */
static inline QE epow(QE q, long double e) {
  const long double partial = q.x * q.x + q.y * q.y + q.z * q.z;
  const long double qMagnitude = sqrtl(partial + q.w * q.w);
  const long double vMagnitude = sqrtl(partial);
  const long double vMagnitudeInv = (vMagnitude > 0) ? 1.0L / vMagnitude : 0;
  const long double scale = e * acosl(q.w / qMagnitude) * vMagnitudeInv;
  const long double magnitude = scale * vMagnitude;
  const long double magnitudeInv = (magnitude > 0) ? 1.0L / magnitude : 0;
  const long double exps = expl(e * logl(qMagnitude));
  const long double scale2 = scale * exps * magnitudeInv * sinl(magnitude);
  QE f = {exps * cosl(magnitude), scale2 * q.x, scale2 * q.y, scale2 * q.z};
  return f;
}
/* QUATERNION_T<Float>::log, QUATERNION_T.h:444-457: the vector part goes
 * through VEC3F (double) and its sqrtf magnitude, acos in double. */
static inline QE elog(QE q) {
  const long double qMagnitude = emag(q);
  const double vx = (double)q.x, vy = (double)q.y, vz = (double)q.z;
  const double vMagnitude = vec3_magnitude(vx, vy, vz);
  double nx = 0, ny = 0, nz = 0;
  if (vMagnitude > 0) { nx = vx / vMagnitude; ny = vy / vMagnitude; nz = vz / vMagnitude; }
  const double a = acos((double)(q.w / (long double)(double)qMagnitude));
  QE f = {logl(qMagnitude), (long double)(nx * a), (long double)(ny * a), (long double)(nz * a)};
  return f;
}

/* ---------------- polynomials ---------------- */

static inline Q root_of(const qo_poly* p, int i) {
  Q r = {p->roots[4 * i + 0], p->roots[4 * i + 1], p->roots[4 * i + 2], p->roots[4 * i + 3]};
  return r;
}

/* POLYNOMIAL_4D::evaluateScaledPowerFactored(point), POLYNOMIAL_4D.cpp:912-923:
 * ordered left-to-right product of (q - r_i)^(s * p_i). */
static Q poly_eval(const qo_poly* p, Q q) {
  Q result = qpow(qsub(q, root_of(p, 0)), p->power_scalar * p->powers[0]);
  for (int i = 1; i < p->n; i++) result = qmul(result, qpow(qsub(q, root_of(p, i)), p->power_scalar * p->powers[i]));
  return result;
}

/* the cached variant (POLYNOMIAL_4D.cpp:1044-1075): forward prefix products
 * and backward suffix products; fwd/bwd hold n entries each. */
static Q poly_eval_cached(const qo_poly* p, Q q, Q* fwd, Q* bwd, Q* pw) {
  const int n = p->n;
  Q result = qpow(qsub(q, root_of(p, 0)), p->power_scalar * p->powers[0]);
  pw[0] = result;
  fwd[0] = result;
  for (int i = 1; i < n; i++) {
    Q term = qpow(qsub(q, root_of(p, i)), p->power_scalar * p->powers[i]);
    pw[i] = term;
    result = qmul(result, term);
    fwd[i] = result;
  }
  Q current = pw[n - 1];
  bwd[n - 1] = current;
  for (int i = n - 2; i >= 0; i--) {
    current = qmul(pw[i], current);
    bwd[i] = current;
  }
  return result;
}

static QE poly_eval_cached_e(const qo_poly* p, QE q, QE* fwd, QE* bwd, QE* pw) {
  const int n = p->n;
  const long double s = (long double)p->power_scalar;
  QE r0 = {p->roots[0], p->roots[1], p->roots[2], p->roots[3]};
  QE result = epow(esub(q, r0), s * (long double)p->powers[0]);
  pw[0] = result;
  fwd[0] = result;
  for (int i = 1; i < n; i++) {
    QE ri = {p->roots[4 * i], p->roots[4 * i + 1], p->roots[4 * i + 2], p->roots[4 * i + 3]};
    QE term = epow(esub(q, ri), s * (long double)p->powers[i]);
    pw[i] = term;
    result = emul(result, term);
    fwd[i] = result;
  }
  QE current = pw[n - 1];
  bwd[n - 1] = current;
  for (int i = n - 2; i >= 0; i--) {
    current = emul(pw[i], current);
    bwd[i] = current;
  }
  return result;
}

/* ---------------- grid ---------------- */

/* FIELD_3D::cellCenter, FIELD_3D.cpp:818-841, with QUATERNION::toRotationMatrix
 * (QUATERNION.cpp:395-412) and MATRIX3 * VEC3 (MATRIX3.h:200-201). */
void qo_cell_center(const qo_grid* g, int x, int y, int z, double* out) {
  const double dx = g->lengths[0] / g->res[0];
  const double dy = g->lengths[1] / g->res[1];
  const double dz = g->lengths[2] / g->res[2];
  double f0 = g->center[0] - 0.5 * g->lengths[0];
  double f1 = g->center[1] - 0.5 * g->lengths[1];
  double f2 = g->center[2] - 0.5 * g->lengths[2];
  f0 += x * dx;
  f1 += y * dy;
  f2 += z * dz;
  f0 += dx * 0.5;
  f1 += dy * 0.5;
  f2 += dz * 0.5;
  f0 -= g->center[0];
  f1 -= g->center[1];
  f2 -= g->center[2];
  const double w = g->rotation[0], qx = g->rotation[1], qy = g->rotation[2], qz = g->rotation[3];
  const double m00 = 1 - 2 * qy * qy - 2 * qz * qz, m01 = 2 * qx * qy - 2 * w * qz, m02 = 2 * qx * qz + 2 * w * qy;
  const double m10 = 2 * qx * qy + 2 * w * qz, m11 = 1 - 2 * qx * qx - 2 * qz * qz, m12 = 2 * qy * qz - 2 * w * qx;
  const double m20 = 2 * qx * qz - 2 * w * qy, m21 = 2 * qy * qz + 2 * w * qx, m22 = 1 - 2 * qx * qx - 2 * qy * qy;
  const double r0 = m00 * f0 + m01 * f1 + m02 * f2;
  const double r1 = m10 * f0 + m11 * f1 + m12 * f2;
  const double r2 = m20 * f0 + m21 * f1 + m22 * f2;
  out[0] = r0 + g->center[0];
  out[1] = r1 + g->center[1];
  out[2] = r2 + g->center[2];
}

/* FIELD_3D::operator()(VEC3F), FIELD_3D.cpp:966-1029 (float weights) */
double qo_field_lookup(const qo_grid* g, const double* data, const double* position) {
  const int xRes = g->res[0], yRes = g->res[1], zRes = g->res[2];
  const double dx = g->lengths[0] / xRes, dy = g->lengths[1] / yRes, dz = g->lengths[2] / zRes;
  const double invDx = 1.0 / dx, invDy = 1.0 / dy, invDz = 1.0 / dz;
  double c0 = g->center[0] - 0.5 * g->lengths[0];
  double c1 = g->center[1] - 0.5 * g->lengths[1];
  double c2 = g->center[2] - 0.5 * g->lengths[2];
  c0 += 0.5 * dx;
  c1 += 0.5 * dy;
  c2 += 0.5 * dz;
  double p0 = position[0] - c0, p1 = position[1] - c1, p2 = position[2] - c2;
  p0 *= invDx;
  p1 *= invDy;
  p2 *= invDz;
  int x0 = (int)p0, y0 = (int)p1, z0 = (int)p2;
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0; z0 = z0 < 0 ? 0 : z0;
  x1 = x1 < 0 ? 0 : x1; y1 = y1 < 0 ? 0 : y1; z1 = z1 < 0 ? 0 : z1;
  x0 = x0 > xRes - 1 ? xRes - 1 : x0; y0 = y0 > yRes - 1 ? yRes - 1 : y0; z0 = z0 > zRes - 1 ? zRes - 1 : z0;
  x1 = x1 > xRes - 1 ? xRes - 1 : x1; y1 = y1 > yRes - 1 ? yRes - 1 : y1; z1 = z1 > zRes - 1 ? zRes - 1 : z1;
  const float s1 = (float)(p0 - x0), s0 = 1.0f - s1;
  const float t1 = (float)(p1 - y0), t0 = 1.0f - t1;
  const float u1 = (float)(p2 - z0), u0 = 1.0f - u1;
  const int64_t slab = (int64_t)xRes * yRes;
  const double d000 = data[x0 + y0 * xRes + z0 * slab], d010 = data[x0 + y1 * xRes + z0 * slab];
  const double d100 = data[x1 + y0 * xRes + z0 * slab], d110 = data[x1 + y1 * xRes + z0 * slab];
  const double d001 = data[x0 + y0 * xRes + z1 * slab], d011 = data[x0 + y1 * xRes + z1 * slab];
  const double d101 = data[x1 + y0 * xRes + z1 * slab], d111 = data[x1 + y1 * xRes + z1 * slab];
  return u0 * (s0 * (t0 * d000 + t1 * d010) + s1 * (t0 * d100 + t1 * d110)) +
         u1 * (s0 * (t0 * d001 + t1 * d011) + s1 * (t0 * d101 + t1 * d111));
}

/* ---------------- the map ---------------- */

/* RATIONAL_4D::evaluateLagrange (RATIONAL_4D.cpp:27-88): with bottom roots,
 * r = 1, then r = (r * t_i) / b_i for the first min(nt, nb) root pairs, then the
 * remaining top terms multiplied or the remaining bottom terms divided, in
 * order; t_i / b_i are evaluateRoot (POLYNOMIAL_4D.cpp:928-940). Without
 * bottom roots it is the factored top product. */
static Q term_of(const qo_poly* p, int i, Q q) {
  const Q b = {q.w - p->roots[4 * i], q.x - p->roots[4 * i + 1], q.y - p->roots[4 * i + 2],
               q.z - p->roots[4 * i + 3]};
  return qpow(b, p->power_scalar * p->powers[i]);
}

static Q lagrange_eval(const qo_poly* top, const qo_poly* bottom, Q q) {
  if (bottom->n == 0) return poly_eval(top, q);
  Q r = {1.0, 0.0, 0.0, 0.0};
  const int same = top->n > bottom->n ? bottom->n : top->n;
  for (int i = 0; i < same; i++) {
    r = qmul(r, term_of(top, i, q));
    r = qdiv(r, term_of(bottom, i, q));
  }
  for (int i = same; i < top->n; i++) r = qmul(r, term_of(top, i, q));
  for (int i = same; i < bottom->n; i++) r = qdiv(r, term_of(bottom, i, q));
  return r;
}

/* POLYNOMIAL_4D::evaluateScaledPowerFactored (POLYNOMIAL_4D.cpp:912-923) with
 * Real = long double: QUATERNION_T<long double>::pow and operator* */
static QE poly_eval_ld(const qo_poly* p, QE q) {
  const long double s = (long double)p->power_scalar;
  QE r0 = {p->roots[0], p->roots[1], p->roots[2], p->roots[3]};
  QE result = epow(esub(q, r0), s * (long double)p->powers[0]);
  for (int i = 1; i < p->n; i++) {
    QE ri = {p->roots[4 * i], p->roots[4 * i + 1], p->roots[4 * i + 2], p->roots[4 * i + 3]};
    result = emul(result, epow(esub(q, ri), s * (long double)p->powers[i]));
  }
  return result;
}

/* One voxel of computeLogScaledPowerRationalMap (OPTIMIZE_3D.cpp:1317-1362) in
 * the reference's default build, Real = long double (SETTINGS.h:10): every
 * quaternion operation, |q| and the escape / black-hole tests (REAL_MIN stays
 * FLT_MIN, SETTINGS.h:58) in x87 long double; the field value is stored as a
 * double (the FIELD_3D file format, FIELD_3D.cpp:4383-4394). The cell centre
 * is the double one. */
static int iterate_point_ld(Q q0, const qo_poly* top, const qo_poly* bottom, const qo_params* p, double* out,
                            int32_t* iters) {
  QE q = {q0.w, q0.x, q0.y, q0.z};
  long double magnitude = emag(q);
  if ((p->flags & QO_SKIP_ORIGIN) && magnitude < 1e-8L) return 0;
  int it = 0, executed = 0;
  while (magnitude < (long double)p->escape && it < p->max_iter) {
    QE topEval = poly_eval_ld(top, q);
    if (bottom->n > 0) {
      QE bottomEval = poly_eval_ld(bottom, q);
      q = ediv(topEval, bottomEval);
    } else {
      q = topEval;
    }
    q = escale(q, (long double)p->exp_scaling);
    magnitude = emag(q);
    it++;
    executed++;
    if ((p->flags & QO_BLACK_HOLE) && magnitude < 10.0L * (long double)FLT_MIN) it = p->max_iter;
  }
  *out = (double)logl(magnitude);
  if (iters) *iters = executed;
  return 1;
}

/* One voxel of OPTIMIZE_3D::computeLogScaledPowerRationalMap
 * (OPTIMIZE_3D.cpp:1317-1362) / computeRationalMapAP first pass (:1423-1475).
 * Returns 1 when the value was written, 0 for a skipped origin cell. */
static int iterate_point(Q q, const qo_poly* top, const qo_poly* bottom, const qo_params* p,
                         double* out, int32_t* iters, int* overflowed) {
  if (p->flags & QO_EXTENDED) return iterate_point_ld(q, top, bottom, p, out, iters);
  double magnitude = qmag(q);
  if ((p->flags & QO_SKIP_ORIGIN) && magnitude < 1e-8) return 0;
  int it = 0, executed = 0;
  const int ap = (p->flags & QO_AP_REPASS) != 0;
  while (magnitude < p->escape && it < p->max_iter) {
    if (ap) feclearexcept(FE_OVERFLOW);
    if (p->flags & QO_LAGRANGE) {
      q = lagrange_eval(top, bottom, q);  /* computeLagrangeNonlinearSlice, TRIANGLE_MESH.cpp:281-324 */
    } else {
      Q topEval = poly_eval(top, q);
      if (bottom->n > 0) {
        Q bottomEval = poly_eval(bottom, q);
        q = qdiv(topEval, bottomEval);
      } else {
        q = topEval;
      }
    }
    if (ap && fetestexcept(FE_OVERFLOW)) *overflowed = 1;
    q = qscale(q, p->exp_scaling);
    magnitude = qmag(q);
    if (ap && (isinf(magnitude) || isnan(magnitude))) *overflowed = 1;
    it++;
    executed++;
    if ((p->flags & QO_BLACK_HOLE) && magnitude < 10.0 * FLT_MIN) it = p->max_iter;
  }
  *out = log(magnitude);
  if (iters) *iters = executed;
  return 1;
}

/* long double re-pass of one voxel (OPTIMIZE_3D.cpp:1496-1567) */
static int iterate_point_e(const double* c3, double slice, const qo_poly* top, const qo_poly* bottom,
                           const qo_params* p, double* out, int* nan_out) {
  QE q = {c3[0], c3[1], c3[2], slice};
  long double magnitude = emag(q);
  if (magnitude < 1e-8) return 0;
  int it = 0;
  const int nt = top->n, nb = bottom->n;
  QE* buf = (QE*)malloc(sizeof(QE) * 3 * (size_t)(nt > nb ? nt : nb));
  while (magnitude < p->escape && it < p->max_iter) {
    QE topEval = poly_eval_cached_e(top, q, buf, buf + nt, buf + 2 * nt);
    if (nb > 0) {
      QE bottomEval = poly_eval_cached_e(bottom, q, buf, buf + nb, buf + 2 * nb);
      q = ediv(topEval, bottomEval);
    } else {
      q = topEval;
    }
    q = escale(q, (long double)p->exp_scaling);
    magnitude = emag(q);
    it++;
  }
  free(buf);
  double v = (double)logl(magnitude);
  *nan_out = 0;
  if (isnan(v)) { v = 0; *nan_out = 1; }
  *out = v;
  return 1;
}

void qo_field_map(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p,
                  double* field, int32_t* iters, uint8_t* overflowed, int nthreads) {
  qo_field_map_zrange(g, top, bottom, p, 0, g->res[2], field, iters, overflowed, nthreads);
}

/* The same map restricted to z-slabs [z0, z1) of the grid; out arrays hold
 * (z1 - z0) slabs. Used to sample large grids. */
void qo_field_map_zrange(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p, int z0,
                         int z1, double* field, int32_t* iters, uint8_t* overflowed, int nthreads) {
  const int xRes = g->res[0], yRes = g->res[1], zRes = z1 - z0;
  const int64_t slab = (int64_t)xRes * yRes;
  const int ap = (p->flags & QO_AP_REPASS) != 0;
  uint8_t* ovf = overflowed;
  uint8_t* own = NULL;
  if (ap && !ovf) { own = (uint8_t*)calloc((size_t)slab * zRes, 1); ovf = own; }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int y = 0; y < yRes; y++) {
    for (int z = 0; z < zRes; z++)
      for (int x = 0; x < xRes; x++) {
        const int64_t idx = x + y * (int64_t)xRes + z * slab;
        double c[3];
        qo_cell_center(g, x, y, z + z0, c);
        Q q = {c[0], c[1], c[2], p->slice};
        int of = 0;
        int32_t n = 0;
        if (!iterate_point(q, top, bottom, p, &field[idx], &n, &of) && iters) n = 0;
        if (iters) iters[idx] = n;
        if (ovf) ovf[idx] = (uint8_t)of;
      }
  }
  if (ap) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic)
#endif
    for (int y = 0; y < yRes; y++)
      for (int z = 0; z < zRes; z++)
        for (int x = 0; x < xRes; x++) {
          const int64_t idx = x + y * (int64_t)xRes + z * slab;
          if (!ovf[idx]) continue;
          double c[3];
          qo_cell_center(g, x, y, z + z0, c);
          int nan_out = 0;
          iterate_point_e(c, p->slice, top, bottom, p, &field[idx], &nan_out);
        }
  }
  free(own);
}

void qo_points_eval(const double* pts, int64_t n, const qo_poly* top, const qo_poly* bottom,
                    const qo_params* p, double* out, int nthreads) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
#endif
  for (int64_t i = 0; i < n; i++) {
    Q q = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], p->slice};
    int of = 0;
    if (!iterate_point(q, top, bottom, p, &out[i], NULL, &of)) out[i] = NAN;
  }
}

/* ---------------- marching cubes ---------------- */

/* TRIANGLE_MESH::getVertexPair, TRIANGLE_MESH.cpp:1504-1563: reference edge
 * 1..12 -> (offset of the lower endpoint, axis 0=x,1=y,2=z). */
static void edge_of(int e, int64_t xRes, int64_t slab, int64_t* first, int* axis) {
  switch (e) {
    case 1: *first = 0; *axis = 1; break;
    case 2: *first = 0; *axis = 2; break;
    case 3: *first = slab; *axis = 1; break;
    case 4: *first = xRes; *axis = 2; break;
    case 5: *first = 1; *axis = 1; break;
    case 6: *first = 1; *axis = 2; break;
    case 7: *first = 1 + slab; *axis = 1; break;
    case 8: *first = 1 + xRes; *axis = 2; break;
    case 9: *first = xRes; *axis = 0; break;
    case 10: *first = 0; *axis = 0; break;
    case 11: *first = slab; *axis = 0; break;
    default: *first = xRes + slab; *axis = 0; break; /* 12 */
  }
}

static inline int cell_flag(const double* f, int64_t i, int64_t xRes, int64_t slab) {
  /* CUBE holds floats (TRIANGLE_MESH.h:283-288); flag at TRIANGLE_MESH.cpp:1037-1039 */
  const float NNN = (float)f[i], NNP = (float)f[i + slab], NPN = (float)f[i + xRes], NPP = (float)f[i + xRes + slab];
  const float PNN = (float)f[i + 1], PNP = (float)f[i + 1 + slab], PPN = (float)f[i + 1 + xRes],
              PPP = (float)f[i + 1 + xRes + slab];
  return (NNN > 0) + 2 * (NNP > 0) + 4 * (NPN > 0) + 8 * (NPP > 0) + 16 * (PNN > 0) + 32 * (PNP > 0) +
         64 * (PPN > 0) + 128 * (PPP > 0);
}

/* TRIANGLE_MESH::computeStagedMarchingCubes, TRIANGLE_MESH.cpp:989-1114, with
 * computeEdgeInterpolations (:2373-2425) and addVertexPairTriangle (:2490-2503).
 * The reference's std::map<pair<int,int>> orders edges by (first, second);
 * marking referenced edges per lower endpoint and walking endpoints in
 * linear order with axes x,y,z (second = first+1 < first+xRes < first+slab)
 * visits them in the same order. */
int qo_marching_cubes(const qo_grid* g, const double* field, int64_t* n_vertices, int64_t* n_triangles,
                      double* vertices, int32_t* triangles) {
  const int64_t xRes = g->res[0], yRes = g->res[1], zRes = g->res[2];
  const int64_t slab = xRes * yRes, total = slab * zRes;
  uint8_t* mark = (uint8_t*)calloc((size_t)total, 1);
  int32_t* vid = NULL;
  if (!mark) return 1;
  for (int64_t z = 0; z < zRes - 1; z++)
    for (int64_t y = 0; y < yRes - 1; y++)
      for (int64_t x = 0; x < xRes - 1; x++) {
        const int64_t index = x + y * xRes + z * slab;
        const int flag = cell_flag(field, index, xRes, slab);
        for (int k = 0; k < 15 && QO_MC_TRIS[flag][k]; k++) {
          int64_t off;
          int axis;
          edge_of(QO_MC_TRIS[flag][k], xRes, slab, &off, &axis);
          mark[index + off] |= (uint8_t)(1u << axis);
        }
      }
  int64_t nv = 0;
  vid = (int32_t*)malloc(sizeof(int32_t) * (size_t)total);
  if (!vid) { free(mark); return 1; }
  for (int64_t i = 0; i < total; i++) {
    vid[i] = (int32_t)nv;
    nv += (mark[i] & 1) + ((mark[i] >> 1) & 1) + ((mark[i] >> 2) & 1);
  }
  if (vertices) {
    for (int64_t i = 0; i < total; i++) {
      if (!mark[i]) continue;
      int64_t k = vid[i];
      const int x = (int)(i % xRes), y = (int)((i / xRes) % yRes), z = (int)(i / slab);
      for (int axis = 0; axis < 3; axis++) {
        if (!(mark[i] & (1u << axis))) continue;
        const int x2 = x + (axis == 0), y2 = y + (axis == 1), z2 = z + (axis == 2);
        const int64_t j = x2 + y2 * xRes + z2 * slab;
        double a[3], b[3];
        qo_cell_center(g, x, y, z, a);
        qo_cell_center(g, x2, y2, z2, b);
        const double off0 = b[0] - a[0], off1 = b[1] - a[1], off2 = b[2] - a[2];
        const double d0 = field[i], d1 = field[j];
        const double interpolation = d0 / (d0 - d1);
        vertices[3 * k + 0] = a[0] + interpolation * off0;
        vertices[3 * k + 1] = a[1] + interpolation * off1;
        vertices[3 * k + 2] = a[2] + interpolation * off2;
        k++;
      }
    }
  }
  /* triangles in cell order, table order; drop degenerate ones (< 1e-7) */
  int64_t nt = 0;
  for (int64_t z = 0; z < zRes - 1; z++)
    for (int64_t y = 0; y < yRes - 1; y++)
      for (int64_t x = 0; x < xRes - 1; x++) {
        const int64_t index = x + y * xRes + z * slab;
        const int flag = cell_flag(field, index, xRes, slab);
        for (int k = 0; k < 15 && QO_MC_TRIS[flag][k]; k += 3) {
          int32_t v[3];
          for (int c = 0; c < 3; c++) {
            int64_t off;
            int axis;
            edge_of(QO_MC_TRIS[flag][k + c], xRes, slab, &off, &axis);
            const int64_t lo = index + off;
            int32_t id = vid[lo];
            for (int a = 0; a < axis; a++) id += (mark[lo] >> a) & 1;
            v[c] = id;
          }
          if (vertices) {
            const double* p0 = vertices + 3 * (int64_t)v[0];
            const double* p1 = vertices + 3 * (int64_t)v[1];
            const double* p2 = vertices + 3 * (int64_t)v[2];
            double e[3];
            e[0] = p0[0] - p1[0]; e[1] = p0[1] - p1[1]; e[2] = p0[2] - p1[2];
            const double dist0 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            e[0] = p0[0] - p2[0]; e[1] = p0[1] - p2[1]; e[2] = p0[2] - p2[2];
            const double dist1 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            e[0] = p1[0] - p2[0]; e[1] = p1[1] - p2[1]; e[2] = p1[2] - p2[2];
            const double dist2 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            const double eps = 1e-7;
            if (dist0 < eps || dist1 < eps || dist2 < eps) continue;
            if (triangles) {
              triangles[3 * nt + 0] = v[0];
              triangles[3 * nt + 1] = v[1];
              triangles[3 * nt + 2] = v[2];
            }
          }
          nt++;
        }
      }
  *n_vertices = nv;
  *n_triangles = nt; /* without vertices the degenerate filter cannot run: count is pre-filter */
  free(vid);
  free(mark);
  return 0;
}

/* ---------------- non-linear marching cubes ---------------- */

/* TRIANGLE_MESH::nonlinearValue (TRIANGLE_MESH.cpp:1755-1795) */
static double nonlinear_value(const double* pos, const qo_poly* top, const qo_poly* bottom, const qo_params* p,
                              double iso) {
  Q q = {pos[0], pos[1], pos[2], p->slice};
  if (p->flags & QO_LAGRANGE) {
    /* TRIANGLE_MESH::lagrangeNonlinearValue (TRIANGLE_MESH.cpp:1684-1730): an
     * infinite |q| restores the previous iterate and stops. The reference's
     * previous magnitude is uninitialised before the first iteration; here it
     * starts as |q0|. */
    double magnitude = qmag(q), prev = magnitude;
    int it = 0;
    while (magnitude < p->escape && it < p->max_iter) {
      q = qscale(lagrange_eval(top, bottom, q), p->exp_scaling);
      magnitude = qmag(q);
      if (isinf(magnitude)) {
        magnitude = prev;
        break;
      }
      prev = magnitude;
      it++;
      if (magnitude < 10.0 * FLT_MIN) it = p->max_iter;
    }
    /* an infinite final value is returned as FLT_MAX (TRIANGLE_MESH.cpp:1733-1742) */
    const double v = log(magnitude) - iso;
    return isinf(v) ? (double)FLT_MAX : v;
  }
  double F = 0;
  int of = 0;
  iterate_point(q, top, bottom, p, &F, NULL, &of);
  return F - iso;
}

/* TRIANGLE_MESH::midpointSearch (TRIANGLE_MESH.cpp:2232-2271); returns 1 when
 * the end values do not straddle zero (the reference exits there) */
static int midpoint_search(double* pp, double vp, double* pn, double vn, const qo_poly* top, const qo_poly* bottom,
                           const qo_params* p, double iso, double* out) {
  for (int recursion = 0;; recursion++) {
    double m[3] = {pp[0] + pn[0], pp[1] + pn[1], pp[2] + pn[2]};
    m[0] *= 0.5;
    m[1] *= 0.5;
    m[2] *= 0.5;
    out[0] = m[0];
    out[1] = m[1];
    out[2] = m[2];
    if (vp * vn >= 0.0) return 1;
    if (recursion >= 6) return 0;
    const double vm = nonlinear_value(m, top, bottom, p, iso);
    if (fabs(vm) < 1e-8) return 0;
    if (vm < 0) {
      pn[0] = m[0]; pn[1] = m[1]; pn[2] = m[2];
      vn = vm;
    } else {
      pp[0] = m[0]; pp[1] = m[1]; pp[2] = m[2];
      vp = vm;
    }
  }
}

/* TRIANGLE_MESH::computeNonlinearMarchingCubesLowMemoryHuge (TRIANGLE_MESH.cpp:681-768)
 * with computeAllLowMemorySlicesHuge (:4884-4948) and
 * computeNonlinearEdgeInterpolationsHuge (:2034-2123), on the whole field at
 * once (the slab streaming only bounds memory). Cases from the raw field
 * log|q|; vertices by bisection on log|q| - iso. Returns the number of edges
 * whose end values do not straddle zero (the reference exits on the first). */
int64_t qo_nonlinear_mc(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p_in,
                        double iso, int64_t* n_vertices, int64_t* n_triangles, double* vertices, int32_t* triangles) {
  return qo_nonlinear_mc_field(g, NULL, top, bottom, p_in, iso, n_vertices, n_triangles, vertices, triangles);
}

/* The same mesher with the cases taken from a given field (the
 * TRIANGLE_MESH(const FIELD_3D&, top, bottom, ...) constructor,
 * TRIANGLE_MESH.cpp:62-79 -> computeNonlinearMarchingCubes :152-215); with
 * given == NULL the field is the map's own (the slab mesher). */
int64_t qo_nonlinear_mc_field(const qo_grid* g, const double* given, const qo_poly* top, const qo_poly* bottom,
                              const qo_params* p_in, double iso, int64_t* n_vertices, int64_t* n_triangles,
                              double* vertices, int32_t* triangles) {
  /* the mesher's slabs and values always stop at the black hole
   * (TRIANGLE_MESH.cpp:374-376, 1789-1791) */
  qo_params pp = *p_in;
  pp.flags |= QO_BLACK_HOLE;
  const qo_params* p = &pp;
  const int64_t xRes = g->res[0], yRes = g->res[1], zRes = g->res[2];
  const int64_t slab = xRes * yRes, total = slab * zRes;
  double* field = (double*)malloc(sizeof(double) * (size_t)total);
  uint8_t* mark = (uint8_t*)calloc((size_t)total, 1);
  int32_t* vid = (int32_t*)malloc(sizeof(int32_t) * (size_t)total);
  int64_t bad = 0;
  if (given) memcpy(field, given, sizeof(double) * (size_t)total);
  else qo_field_map(g, top, bottom, p, field, NULL, NULL, 0);
  for (int64_t z = 0; z < zRes - 1; z++)
    for (int64_t y = 0; y < yRes - 1; y++)
      for (int64_t x = 0; x < xRes - 1; x++) {
        const int64_t index = x + y * xRes + z * slab;
        const int flag = cell_flag(field, index, xRes, slab);
        for (int k = 0; k < 15 && QO_MC_TRIS[flag][k]; k++) {
          int64_t off;
          int axis;
          edge_of(QO_MC_TRIS[flag][k], xRes, slab, &off, &axis);
          mark[index + off] |= (uint8_t)(1u << axis);
        }
      }
  int64_t nv = 0;
  for (int64_t i = 0; i < total; i++) {
    vid[i] = (int32_t)nv;
    nv += (mark[i] & 1) + ((mark[i] >> 1) & 1) + ((mark[i] >> 2) & 1);
  }
  if (vertices) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
#endif
    for (int64_t i = 0; i < total; i++) {
      if (!mark[i]) continue;
      int64_t k = vid[i];
      const int x = (int)(i % xRes), y = (int)((i / xRes) % yRes), z = (int)(i / slab);
      double a[3];
      qo_cell_center(g, x, y, z, a);
      const double va = nonlinear_value(a, top, bottom, p, iso);
      for (int axis = 0; axis < 3; axis++) {
        if (!(mark[i] & (1u << axis))) continue;
        double b[3];
        qo_cell_center(g, x + (axis == 0), y + (axis == 1), z + (axis == 2), b);
        const double vb = nonlinear_value(b, top, bottom, p, iso);
        double pp[3], pn[3], vp, vn;
        if (va * vb >= 0.0) bad++;
        if (va < 0) {
          memcpy(pp, b, sizeof(pp)); vp = vb; memcpy(pn, a, sizeof(pn)); vn = va;
        } else {
          memcpy(pp, a, sizeof(pp)); vp = va; memcpy(pn, b, sizeof(pn)); vn = vb;
        }
        if (midpoint_search(pp, vp, pn, vn, top, bottom, p, iso, vertices + 3 * k) && !(va * vb >= 0.0)) bad++;
        k++;
      }
    }
  }
  int64_t nt = 0;
  for (int64_t z = 0; z < zRes - 1; z++)
    for (int64_t y = 0; y < yRes - 1; y++)
      for (int64_t x = 0; x < xRes - 1; x++) {
        const int64_t index = x + y * xRes + z * slab;
        const int flag = cell_flag(field, index, xRes, slab);
        for (int k = 0; k < 15 && QO_MC_TRIS[flag][k]; k += 3) {
          int32_t v[3];
          for (int c = 0; c < 3; c++) {
            int64_t off;
            int axis;
            edge_of(QO_MC_TRIS[flag][k + c], xRes, slab, &off, &axis);
            const int64_t lo = index + off;
            int32_t id = vid[lo];
            for (int a2 = 0; a2 < axis; a2++) id += (mark[lo] >> a2) & 1;
            v[c] = id;
          }
          if (vertices) {
            const double* p0 = vertices + 3 * (int64_t)v[0];
            const double* p1 = vertices + 3 * (int64_t)v[1];
            const double* p2 = vertices + 3 * (int64_t)v[2];
            double e[3];
            e[0] = p0[0] - p1[0]; e[1] = p0[1] - p1[1]; e[2] = p0[2] - p1[2];
            const double d0 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            e[0] = p0[0] - p2[0]; e[1] = p0[1] - p2[1]; e[2] = p0[2] - p2[2];
            const double d1 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            e[0] = p1[0] - p2[0]; e[1] = p1[1] - p2[1]; e[2] = p1[2] - p2[2];
            const double d2 = sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
            if (d0 < 1e-7 || d1 < 1e-7 || d2 < 1e-7) continue;
            if (triangles) {
              triangles[3 * nt + 0] = v[0];
              triangles[3 * nt + 1] = v[1];
              triangles[3 * nt + 2] = v[2];
            }
          }
          nt++;
        }
      }
  *n_vertices = nv;
  *n_triangles = nt;
  free(field);
  free(mark);
  free(vid);
  return bad;
}

/* ---------------- FIELD_3D clean-up passes ---------------- */

/* FIELD_3D::clampInfsToMean (FIELD_3D.cpp:5182-5207) and clampNansToMean
 * (:5212-5234): serial sum in index order of the cells that are not infinite
 * (not NaN), mean = sum / count, then the infinite (NaN) cells take the mean.
 * The reference counts in an int; the count here is 64-bit (the same for
 * fields under 2^31 cells). Returns the mean. */
static double clamp_to_mean(double* data, int64_t n, int nans) {
  double sum = 0;
  int64_t seen = 0;
  for (int64_t i = 0; i < n; i++)
    if (nans ? !isnan(data[i]) : !isinf(data[i])) {
      sum += data[i];
      seen++;
    }
  const double mean = sum / (double)seen;
  for (int64_t i = 0; i < n; i++)
    if (nans ? isnan(data[i]) : isinf(data[i])) data[i] = mean;
  return mean;
}
double qo_clamp_infs_to_mean(double* data, int64_t n) { return clamp_to_mean(data, n, 0); }
double qo_clamp_nans_to_mean(double* data, int64_t n) { return clamp_to_mean(data, n, 1); }

/* FIELD_3D::setBorders (FIELD_3D.cpp:5357-5382) */
void qo_set_borders(const qo_grid* g, double* data, double value) {
  const int64_t X = g->res[0], Y = g->res[1], Z = g->res[2];
#define AT(x, y, z) data[(x) + (y) * X + (z) * X * Y]
  for (int64_t x = 0; x < X; x++)
    for (int64_t y = 0; y < Y; y++) {
      AT(x, y, 0) = value;
      AT(x, y, Z - 1) = value;
    }
  for (int64_t x = 0; x < X; x++)
    for (int64_t z = 0; z < Z; z++) {
      AT(x, 0, z) = value;
      AT(x, Y - 1, z) = value;
    }
  for (int64_t y = 0; y < Y; y++)
    for (int64_t z = 0; z < Z; z++) {
      AT(0, y, z) = value;
      AT(X - 1, y, z) = value;
    }
#undef AT
}

/* ---------------- scoring fields of the shape optimizer ---------------- */

/* FIELD_3D::outerShell + innerShell (FIELD_3D.cpp:5070-5145): 1 where a cell
 * of one sign has a 26-neighbour of the other sign (zero cells have none). */
static void shells(const qo_grid* g, const double* d, double* out) {
  const int64_t X = g->res[0], Y = g->res[1], Z = g->res[2], XY = X * Y;
  for (int64_t z = 0; z < Z; z++)
    for (int64_t y = 0; y < Y; y++)
      for (int64_t x = 0; x < X; x++) {
        const double c = d[x + y * X + z * XY];
        int flip = 0;
        for (int i = -1; i <= 1; i++)
          for (int j = -1; j <= 1; j++)
            for (int k = -1; k <= 1; k++) {
              if (i == 0 && j == 0 && k == 0) continue;
              const int64_t xn = x + i, yn = y + j, zn = z + k;
              if (xn < 0 || xn > X - 1 || yn < 0 || yn > Y - 1 || zn < 0 || zn > Z - 1) continue;
              if (d[xn + yn * X + zn * XY] * c < 0) flip = 1;
            }
        /* outer shell for c >= 0, inner shell for c <= 0: one of them, or
         * neither at c == 0 (no product with 0 is negative) */
        out[x + y * X + z * XY] = flip ? 1.0 : 0.0;
      }
}

/* FIELD_3D::invertedDistance (FIELD_3D.cpp:5241-5297): the shell cells get
 * 1, then sweep after sweep (each reading the previous sweep) every cell
 * still at or below 1e-5 takes 1 + the smallest of its 6-neighbours above
 * 1e-5; then 1 / value, negated inside (distance < 0) and quartered outside.
 * The reference sweeps forever when a cell can never be reached; here that
 * returns -1. */
int qo_inverted_distance(const qo_grid* g, const double* dist, double* out) {
  const int64_t X = g->res[0], Y = g->res[1], Z = g->res[2], XY = X * Y, n = XY * Z;
  double* cur = (double*)malloc(sizeof(double) * (size_t)n);
  shells(g, dist, out);
  const double eps = 1e-5;
  int status = 0;
  for (;;) {
    int64_t zeros = 0, updated = 0;
    memcpy(cur, out, sizeof(double) * (size_t)n);
    for (int64_t z = 0; z < Z; z++)
      for (int64_t y = 0; y < Y; y++)
        for (int64_t x = 0; x < X; x++) {
          const int64_t i = x + y * X + z * XY;
          if (out[i] > eps) continue;
          zeros++;
          double nb[6];
          nb[0] = (x != 0) ? out[i - 1] : 0;
          nb[1] = (x != X - 1) ? out[i + 1] : 0;
          nb[2] = (y != 0) ? out[i - X] : 0;
          nb[3] = (y != Y - 1) ? out[i + X] : 0;
          nb[4] = (z != 0) ? out[i - XY] : 0;
          nb[5] = (z != Z - 1) ? out[i + XY] : 0;
          double biggest = (double)n + 1;
          for (int k = 0; k < 6; k++)
            if (nb[k] > 1e-5 && nb[k] < biggest) biggest = nb[k];
          if (biggest < (double)n) {
            cur[i] = biggest + 1;
            updated++;
          }
        }
    memcpy(out, cur, sizeof(double) * (size_t)n);
    if (zeros == 0) break;
    if (updated == 0) {
      status = -1;
      break;
    }
  }
  free(cur);
  for (int64_t i = 0; i < n; i++) {
    out[i] = 1.0 / out[i];
    if (dist[i] < 0) out[i] *= -1;
    else out[i] *= 0.25;
  }
  return status;
}

/* FIELD_3D::invertedDistanceNormalized (FIELD_3D.cpp:5316-5352): the field
 * over the sum of its absolute values, then the inside and the outside each
 * scaled to sum to 0.5 in absolute value; *abs_sum receives the absolute sum
 * of the result (OPTIMIZE_3D::read's _inverseFieldSum, OPTIMIZE_3D.cpp:3375-3378).
 * Serial sums in index order. */
int qo_inverted_distance_normalized(const qo_grid* g, const double* dist, double* out, double* abs_sum) {
  const int64_t n = (int64_t)g->res[0] * g->res[1] * g->res[2];
  const int status = qo_inverted_distance(g, dist, out);
  double t = 0;
  for (int64_t i = 0; i < n; i++) t += fabs(out[i]);
  const double scale = 1.0 / t;
  for (int64_t i = 0; i < n; i++) out[i] *= scale;
  double inside = 0, outside = 0;
  for (int64_t i = 0; i < n; i++) {
    if (dist[i] < 0) inside += out[i];
    else outside += out[i];
  }
  const double in_n = 0.5 / fabs(inside), out_n = 0.5 / fabs(outside);
  for (int64_t i = 0; i < n; i++) out[i] *= (dist[i] < 0) ? in_n : out_n;
  if (abs_sum) {
    double a = 0;
    for (int64_t i = 0; i < n; i++) a += fabs(out[i]);
    *abs_sum = a;
  }
  return status;
}

/* OPTIMIZE_3D::weightScore (OPTIMIZE_3D.cpp:4502-4517) with the threshold of
 * inverseDistanceScore (:4522-4563; 0 in weightScore):
 * sum_cells -tanh(alpha (F - threshold)) w(cell centre), and
 * computeMaximumScaledPowerRationalScore (:4484-4498): sum |w(cell centre)|.
 * w is the trilinear lookup of the weight field (FIELD_3D.cpp:966-1029). */
void qo_weight_score(const qo_grid* g, const double* field, const qo_grid* wg, const double* weight, double alpha,
                     double threshold, double* score, double* max_score) {
  const int64_t X = g->res[0], Y = g->res[1], Z = g->res[2];
  double s = 0, m = 0;
  for (int64_t z = 0; z < Z; z++)
    for (int64_t y = 0; y < Y; y++)
      for (int64_t x = 0; x < X; x++) {
        double c[3];
        qo_cell_center(g, (int)x, (int)y, (int)z, c);
        const double w = qo_field_lookup(wg, weight, c);
        s += -tanh(alpha * (field[x + y * X + z * X * Y] - threshold)) * w;
        m += fabs(w);
      }
  if (score) *score = s;
  if (max_score) *max_score = m;
}

/* ---------------- AP score and gradients at one iteration ---------------- */

typedef struct {
  double F;
  Q R, topEval, bottomEval;
  double weight;
} cell_state;

/* d/dp_i of the tanh score for one cell and one top root
 * (computeTopGradientAP body, OPTIMIZE_3D.cpp:8942-9000); bulk != 0 uses the
 * bulk-scalar middle term (computeTopBulkScalarGradientAP, :9310-9317). */
static double top_term(const qo_poly* top, int i, Q q0, const Q* fwd, const Q* bwd, Q R, Q bottomEval,
                       double dSdF, int bulk) {
  Q left = {1, 0, 0, 0}, right = {1, 0, 0, 0};
  if (top->n > 1) {
    if (i > 0) left = fwd[i - 1];
    if (i < top->n - 1) right = bwd[i + 1];
  }
  const double s = top->power_scalar;
  Q base = qsub(q0, root_of(top, i));
  Q power = qpow(base, s * top->powers[i]);
  Q middle = qmul(qscale(power, bulk ? top->powers[i] : s), qlog(base));
  Q topDerivative = qmul(qmul(left, middle), right);
  Q Rderivative = qdiv(topDerivative, bottomEval);
  double dFdt = (1.0 / qdot(R, R)) * qdot(R, Rderivative);
  return dSdF * dFdt;
}

/* computeBottomGradientAP body (OPTIMIZE_3D.cpp:10062-10087) and the bulk
 * variant (computeBottomBulkScalarGradientAP, :10422-10430). */
static double bottom_term(const qo_poly* bottom, int i, Q q0, const Q* fwd, const Q* bwd, Q R, Q topEval,
                          double dSdF, int bulk) {
  Q left = {1, 0, 0, 0}, right = {1, 0, 0, 0};
  if (bottom->n > 1) {
    if (i > 0) left = fwd[i - 1];
    if (i < bottom->n - 1) right = bwd[i + 1];
  }
  const double s = bottom->power_scalar;
  Q base = qsub(q0, root_of(bottom, i));
  Q power = qpow(base, s * bottom->powers[i]);
  Q middle = qmul(qscale(power, bulk ? 1.0 / bottom->powers[i] : 1.0 / s), qinv(qlog(base)));
  Q m1 = {-1, 0, 0, 0};
  Q bottomDerivative = qmul(qmul(qmul(m1, left), middle), right);
  Q Rderivative = qdiv(topEval, bottomDerivative);
  double dFdt = (1.0 / qdot(R, R)) * qdot(R, Rderivative);
  return dSdF * dFdt;
}

/* long double twins for the adaptive pass (OPTIMIZE_3D.cpp:9036-9103) */
static QE q2e(Q q) { QE e = {q.w, q.x, q.y, q.z}; return e; }
static long double top_term_e(const qo_poly* top, int i, QE q0, const QE* fwd, const QE* bwd, QE R, QE bottomEval,
                              long double dSdF, int bulk) {
  QE left = {1, 0, 0, 0}, right = {1, 0, 0, 0};
  if (top->n > 1) {
    if (i > 0) left = fwd[i - 1];
    if (i < top->n - 1) right = bwd[i + 1];
  }
  const long double s = top->power_scalar;
  QE base = esub(q0, q2e(root_of(top, i)));
  QE power = epow(base, s * (long double)top->powers[i]);
  QE middle = emul(escale(power, bulk ? (long double)top->powers[i] : s), elog(base));
  QE topDerivative = emul(emul(left, middle), right);
  QE Rderivative = ediv(topDerivative, bottomEval);
  long double dFdt = (1.0L / edot(R, R)) * edot(R, Rderivative);
  return dSdF * dFdt;
}
static long double bottom_term_e(const qo_poly* bottom, int i, QE q0, const QE* fwd, const QE* bwd, QE R,
                                 QE topEval, long double dSdF, int bulk) {
  QE left = {1, 0, 0, 0}, right = {1, 0, 0, 0};
  if (bottom->n > 1) {
    if (i > 0) left = fwd[i - 1];
    if (i < bottom->n - 1) right = bwd[i + 1];
  }
  const long double s = bottom->power_scalar;
  QE base = esub(q0, q2e(root_of(bottom, i)));
  QE power = epow(base, s * (long double)bottom->powers[i]);
  QE middle = emul(escale(power, bulk ? 1.0L / (long double)bottom->powers[i] : 1.0L / s), einv(elog(base)));
  QE m1 = {-1, 0, 0, 0};
  QE bottomDerivative = emul(emul(emul(m1, left), middle), right);
  QE Rderivative = ediv(topEval, bottomDerivative);
  long double dFdt = (1.0L / edot(R, R)) * edot(R, Rderivative);
  return dSdF * dFdt;
}

/*
 * One iteration of the AP map (computeRationalMapAP with maxIterations 1,
 * OPTIMIZE_3D.cpp:1374-1568), then per-cell: weightScore (:4502-4517),
 * computeExpGradientAP (:3589-3625), compute{Top,Bottom}GradientAP
 * (:8905-9106, :10012-10200) and the bulk-scalar gradients (:9250, :10360).
 *
 * Deviations, each in an error path only: (1) the reference keeps the NaN/inf
 * flags in fields shared by all roots that OpenMP evaluates concurrently, so a
 * cell flagged by one root is re-added by others; here each (cell, root) term
 * that is non-finite in FP64 is recomputed once in long double. (2) The
 * reference's bottom bulk-scalar re-pass zeroes the per-slab sums first
 * (:10456); here the re-pass only replaces the non-finite terms. (3) A cell
 * whose |q0| >= escape never fills its cache, which the reference then reads
 * (UB); such cells are skipped.
 */
void qo_ap_score_gradients(const qo_grid* g, const qo_poly* top, const qo_poly* bottom, const qo_params* p,
                           const qo_grid* wg, const double* weight, double tanh_alpha, double tanh_threshold,
                           double* field, double* grad_top, double* grad_bottom, qo_ap_result* res, int nthreads) {
  const int xRes = g->res[0], yRes = g->res[1], zRes = g->res[2];
  const int64_t slab = (int64_t)xRes * yRes;
  const int nt = top->n, nb = bottom->n;
  double score = 0, gexp = 0, gts = 0, gbs = 0;
  int64_t novf = 0;
  for (int i = 0; i < nt; i++) grad_top[i] = 0;
  for (int i = 0; i < nb; i++) grad_bottom[i] = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  const int nz = zRes;
  double* zsum = (double*)calloc((size_t)nz * (4 + nt + nb), sizeof(double));
  int64_t* zovf = (int64_t*)calloc((size_t)nz, sizeof(int64_t));
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic)
#endif
  for (int z = 0; z < zRes; z++) {
    double* acc = zsum + (size_t)z * (4 + nt + nb);
    const int mx = nt > nb ? nt : nb;
    Q* tf = (Q*)malloc(sizeof(Q) * 3 * (size_t)(nt > 0 ? nt : 1));
    Q* bf = (Q*)malloc(sizeof(Q) * 3 * (size_t)(nb > 0 ? nb : 1));
    QE* te = (QE*)malloc(sizeof(QE) * 3 * (size_t)(mx > 0 ? mx : 1));
    QE* be = (QE*)malloc(sizeof(QE) * 3 * (size_t)(mx > 0 ? mx : 1));
    for (int y = 0; y < yRes; y++)
      for (int x = 0; x < xRes; x++) {
        const int64_t idx = x + y * (int64_t)xRes + z * slab;
        double c[3];
        qo_cell_center(g, x, y, z, c);
        Q q0 = {c[0], c[1], c[2], p->slice};
        const double w = qo_field_lookup(wg, weight, c);
        double mag0 = qmag(q0);
        if (mag0 < 1e-8) {
          /* unwritten cell: the score and exp gradient still read the field */
          const double F = field[idx];
          acc[0] += -tanh(tanh_alpha * F) * w;
          const double S = tanh(tanh_alpha * (F - tanh_threshold));
          acc[1] += -tanh_alpha * w * (1.0 - S * S);
          continue;
        }
        if (!(mag0 < p->escape) || p->max_iter < 1) {
          field[idx] = log(mag0);
          const double F = field[idx];
          acc[0] += -tanh(tanh_alpha * F) * w;
          const double S = tanh(tanh_alpha * (F - tanh_threshold));
          acc[1] += -tanh_alpha * w * (1.0 - S * S);
          continue;
        }
        /* AP first pass, one iteration */
        feclearexcept(FE_OVERFLOW);
        Q topEval = poly_eval_cached(top, q0, tf, tf + nt, tf + 2 * nt);
        Q bottomEval = {1, 0, 0, 0};
        Q qn;
        if (nb > 0) {
          bottomEval = poly_eval_cached(bottom, q0, bf, bf + nb, bf + 2 * nb);
          qn = qdiv(topEval, bottomEval);
        } else {
          qn = topEval;
        }
        int ovf = fetestexcept(FE_OVERFLOW) != 0;
        qn = qscale(qn, p->exp_scaling);
        double magnitude = qmag(qn);
        if (isinf(magnitude) || isnan(magnitude)) ovf = 1;
        double F = log(magnitude);
        QE q0e = {c[0], c[1], c[2], p->slice};
        QE topE = {1, 0, 0, 0}, botE = {1, 0, 0, 0}, RE;
        int have_e = 0;
        if (ovf) {
          topE = poly_eval_cached_e(top, q0e, te, te + nt, te + 2 * nt);
          if (nb > 0) {
            botE = poly_eval_cached_e(bottom, q0e, be, be + nb, be + 2 * nb);
            RE = ediv(topE, botE);
          } else {
            RE = topE;
          }
          long double me = emag(escale(RE, (long double)p->exp_scaling));
          double v = (double)logl(me);
          if (isnan(v)) v = 0;
          F = v;
          have_e = 1;
          zovf[z]++;
        }
        field[idx] = F;
        acc[0] += -tanh(tanh_alpha * F) * w;
        {
          const double S = tanh(tanh_alpha * (F - tanh_threshold));
          acc[1] += -tanh_alpha * w * (1.0 - S * S);
        }
        /* gradients */
        Q R = qdiv(topEval, bottomEval);
        double Rmagnitude = qmag(R);
        /* a non-overflowed cell whose |R| is 0 or inf is skipped (OPTIMIZE_3D.cpp:8956);
         * overflowed cells are redone in long double; in FP64 cells a non-finite
         * term is redone in long double */
        const int skip_cell = !ovf && (Rmagnitude <= 0.0 || isinf(Rmagnitude) || isnan(Rmagnitude));
        int fp64_ok = !ovf && !skip_cell;
        double Fg = log(Rmagnitude) + log(p->exp_scaling);
        double tanhTerm = tanh(tanh_alpha * (Fg - tanh_threshold));
        double dSdF = -tanh_alpha * w * (1 - tanhTerm * tanhTerm);
        long double dSdFe = 0;
        QE Re = {0, 0, 0, 0};
        int e_ok = 0;
        #define ENSURE_E()                                                                 \
          do {                                                                             \
            if (!have_e) {                                                                 \
              topE = poly_eval_cached_e(top, q0e, te, te + nt, te + 2 * nt);               \
              if (nb > 0) botE = poly_eval_cached_e(bottom, q0e, be, be + nb, be + 2 * nb); \
              have_e = 1;                                                                  \
            }                                                                              \
            if (!e_ok) {                                                                   \
              Re = ediv(topE, botE);                                                       \
              long double Rm = emag(Re);                                                   \
              e_ok = (Rm <= 0.0L || isinf(Rm)) ? -1 : 1;                                   \
              long double Fe = logl(Rm) + logl((long double)p->exp_scaling);               \
              long double tt = tanhl((long double)tanh_alpha * (Fe - tanh_threshold));     \
              dSdFe = -(long double)tanh_alpha * w * (1 - tt * tt);                        \
            }                                                                              \
          } while (0)
        if (skip_cell) continue;
        if (!fp64_ok) ENSURE_E();
        for (int i = 0; i < nt; i++) {
          for (int bulk = 0; bulk < 2; bulk++) {
            double v = 0;
            int done = 0;
            if (fp64_ok) {
              v = top_term(top, i, q0, tf, tf + nt, R, bottomEval, dSdF, bulk);
              done = !(isnan(v) || isinf(v));
            }
            if (!done) {
              ENSURE_E();
              v = (e_ok > 0) ? (double)top_term_e(top, i, q0e, te, te + nt, Re, botE, dSdFe, bulk) : 0.0;
            }
            acc[bulk ? 2 : 4 + i] += v;
          }
        }
        if (nb > 0 && bottom->power_scalar > 0.0) {
          for (int i = 0; i < nb; i++) {
            double v = 0;
            int done = 0;
            if (fp64_ok) {
              v = bottom_term(bottom, i, q0, bf, bf + nb, R, topEval, dSdF, 0);
              done = !(isnan(v) || isinf(v));
            }
            if (!done) {
              ENSURE_E();
              v = (e_ok > 0) ? (double)bottom_term_e(bottom, i, q0e, be, be + nb, Re, topE, dSdFe, 0) : 0.0;
            }
            acc[4 + nt + i] += v;
          }
        }
        if (nb > 0) {
          for (int i = 0; i < nb; i++) {
            if (bottom->powers[i] < 1e-8) continue;
            double v = 0;
            int done = 0;
            if (fp64_ok) {
              v = bottom_term(bottom, i, q0, bf, bf + nb, R, topEval, dSdF, 1);
              done = !(isnan(v) || isinf(v));
            }
            if (!done) {
              ENSURE_E();
              v = (e_ok > 0) ? (double)bottom_term_e(bottom, i, q0e, be, be + nb, Re, topE, dSdFe, 1) : 0.0;
            }
            acc[3] += v;
          }
        }
        #undef ENSURE_E
      }
    free(tf);
    free(bf);
    free(te);
    free(be);
  }
  for (int z = 0; z < nz; z++) {
    const double* acc = zsum + (size_t)z * (4 + nt + nb);
    score += acc[0];
    gexp += acc[1];
    gts += acc[2];
    gbs += acc[3];
    for (int i = 0; i < nt; i++) grad_top[i] += acc[4 + i];
    for (int i = 0; i < nb; i++) grad_bottom[i] += acc[4 + nt + i];
    novf += zovf[z];
  }
  free(zsum);
  free(zovf);
  res->score = score;
  res->grad_exp = gexp;
  res->grad_top_scalar = gts;
  res->grad_bottom_scalar = gbs;
  res->overflowed = novf;
}

/* ---------------- exported primitives for the algebra tests ---------------- */

static inline Q load_q(const double* a) { Q q = {a[0], a[1], a[2], a[3]}; return q; }
static inline void store_q(Q q, double* a) { a[0] = q.w; a[1] = q.x; a[2] = q.y; a[3] = q.z; }
void qo_qmul(const double* l, const double* r, double* out) { store_q(qmul(load_q(l), load_q(r)), out); }
void qo_qdiv(const double* l, const double* r, double* out) { store_q(qdiv(load_q(l), load_q(r)), out); }
void qo_qpow(const double* q, double e, double* out) { store_q(qpow(load_q(q), e), out); }
void qo_qlog(const double* q, double* out) { store_q(qlog(load_q(q)), out); }
double qo_qmag(const double* q) { return qmag(load_q(q)); }
void qo_poly_eval(const qo_poly* p, const double* q, double* out) { store_q(poly_eval(p, load_q(q)), out); }
