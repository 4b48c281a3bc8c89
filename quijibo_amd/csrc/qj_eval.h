// FP64 evaluation of the map at one point, shared by the field kernels and
// the non-linear mesher (OPTIMIZE_3D.cpp:1317-1362, TRIANGLE_MESH.cpp:1755-1795).
#pragma once
#include <float.h>

#include "qj_internal.h"
#include "qj_quat.h"

namespace qj {

constexpr double kBlackHole = 10.0 * (double)FLT_MIN;  // 10 * REAL_MIN, SETTINGS.h:58, OPTIMIZE_3D.cpp:1358

// ---------------------------------------------------------------- FP64 method

// POLYNOMIAL_4D::evaluateScaledPowerFactored (POLYNOMIAL_4D.cpp:912-923)
// Once a component of the running product is NaN every later product is NaN
// in all four components (each output component reads all four inputs), and
// so are T B^-1 and |q|: the loop can stop there with the same F.
__device__ __forceinline__ bool has_nan(const q64& q) { return isnan(q.w) || isnan(q.x) || isnan(q.y) || isnan(q.z); }

// How term64 takes (q - r)^e:
//   kPolar      QUATERNION::pow's polar form for every exponent (QUATERNION.cpp:375-393)
//   kExactInt   exponents exactly 1 and 2 as b and b*b, others polar
//   kExactOnly  every exponent is 1 or 2 (the caller checked), no polar code
// The exact forms agree with the polar form to a few ulps and keep its special
// cases: b = 0 gives NaN, v = 0 gives (|w|^e, +0 * v).
enum PowMode { kPolar = 0, kExactInt = 1, kExactOnly = 2 };

// One root record by value through any pointer type (generic, or the constant
// address space where a kernel wants scalar loads of a pointer it read from
// memory).
template <class RP>
__device__ __forceinline__ RootRec64 load_root64(RP p) {
  return RootRec64{p->w, p->x, p->y, p->z, p->e, p->p};
}

template <int kPow>
__device__ __forceinline__ q64 term64(const RootRec64& r, const q64& q) {
  const q64 b = qsub(q, q64{r.w, r.x, r.y, r.z});
  if (kPow == kExactOnly || (kPow == kExactInt && (r.e == 1.0 || r.e == 2.0))) {
    const double partial = b.x * b.x + b.y * b.y + b.z * b.z;
    if (!(partial + b.w * b.w > 0.0)) {
      const double nan = __builtin_nan("");
      return q64{nan, nan, nan, nan};
    }
    if (!(partial > 0.0)) {
      const double m = r.e == 1.0 ? fabs(b.w) : b.w * b.w;
      return q64{m, 0.0 * b.x, 0.0 * b.y, 0.0 * b.z};
    }
    if (r.e == 1.0) return b;
    const double t = 2.0 * b.w;
    return q64{b.w * b.w - partial, t * b.x, t * b.y, t * b.z};
  }
  if (kPow == kExactOnly) return b;  // not reached: kExactOnly callers have only e = 1, 2
  return qpow(b, r.e);
}

template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ q64 poly64(RP r, int n, const q64& q) {
  q64 res = term64<kPow>(load_root64(r), q);
  for (int i = 1; i < n; i++) {
    res = qmul(res, term64<kPow>(load_root64(r + i), q));
    if ((i & 7) == 0 && has_nan(res)) {
      const double nan = __builtin_nan("");
      return q64{nan, nan, nan, nan};
    }
  }
  return res;
}

__device__ __forceinline__ bool finite4(const q64& q) {
  return isfinite(q.w) && isfinite(q.x) && isfinite(q.y) && isfinite(q.z);
}

struct Res64 {
  double F;
  int iters;
  bool written;
  bool overflow;
};

// RATIONAL_4D::evaluateLagrange (RATIONAL_4D.cpp:27-88): with bottom roots,
// r = 1, then r = (r t_i) / b_i for the paired roots, then the leftover top
// terms multiplied or bottom terms divided; the factored top product otherwise.
template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ q64 lagrange64(RP roots, int nt, int nb, const q64& q) {
  if (nb == 0) return poly64<kPow>(roots, nt, q);
  const RP bot = roots + nt;
  q64 r{1.0, 0.0, 0.0, 0.0};
  const int same = nt < nb ? nt : nb;
  for (int i = 0; i < same; i++) {
    r = qmul(r, term64<kPow>(load_root64(roots + i), q));
    r = qdiv(r, term64<kPow>(load_root64(bot + i), q));
  }
  for (int i = same; i < nt; i++) r = qmul(r, term64<kPow>(load_root64(roots + i), q));
  for (int i = same; i < nb; i++) r = qdiv(r, term64<kPow>(load_root64(bot + i), q));
  return r;
}

// One voxel of computeLogScaledPowerRationalMap (OPTIMIZE_3D.cpp:1317-1362) /
// computeRationalMapAP's first pass (:1423-1475). Overflow detection replaces
// fetestexcept(FE_OVERFLOW): any non-finite T, B, |B|^2, quotient or |q|.
template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ Res64 eval64(const MapArgs& a, RP roots, q64 q) {
  Res64 r{0.0, 0, true, false};
  double magnitude = qmag(q);
  if ((a.flags & QJ_MAP_SKIP_ORIGIN) && magnitude < 1e-8) {
    r.written = false;
    return r;
  }
  const bool adaptive = (a.flags & QJ_MAP_ADAPTIVE) != 0;
  int it = 0;
  while (magnitude < a.escape && it < a.max_iter) {
    if (a.flags & QJ_MAP_LAGRANGE) {
      q = qscale(lagrange64<kPow>(roots, a.nt, a.nb, q), a.exp_scaling);
      magnitude = qmag(q);
      it++;
      r.iters++;
      if ((a.flags & QJ_MAP_BLACK_HOLE) && magnitude < kBlackHole) it = a.max_iter;
      continue;
    }
    const q64 T = poly64<kPow>(roots, a.nt, q);
    if (a.nb > 0) {
      if (has_nan(T) && !adaptive) {
        // T B^-1 is all-NaN whatever B is; skip evaluating B
        const double nan = __builtin_nan("");
        q = q64{nan, nan, nan, nan};
        magnitude = nan;
        it++;
        r.iters++;
        break;
      }
      const q64 B = poly64<kPow>(roots + a.nt, a.nb, q);
      q = qdiv(T, B);
      if (adaptive) {
        const double m2 = B.w * B.w + B.x * B.x + B.y * B.y + B.z * B.z;
        r.overflow |= !(finite4(T) && finite4(B) && isfinite(m2) && finite4(q));
      }
    } else {
      q = T;
      if (adaptive) r.overflow |= !finite4(T);
    }
    q = qscale(q, a.exp_scaling);
    magnitude = qmag(q);
    if (adaptive && !isfinite(magnitude)) r.overflow = true;
    it++;
    r.iters++;
    if ((a.flags & QJ_MAP_BLACK_HOLE) && magnitude < kBlackHole) it = a.max_iter;
  }
  r.F = log(magnitude);
  return r;
}


}  // namespace qj
