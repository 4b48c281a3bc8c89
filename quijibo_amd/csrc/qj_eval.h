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

// The exact forms' arithmetic with every rounding fixed by the source
// (explicit fused multiply-adds, contraction off), so that the fast and the
// guarded product below, and every kernel that inlines them, round alike.
__device__ __forceinline__ q64 qsub64x(const q64& a, const RootRec64& r) {
#pragma clang fp contract(off)
  return q64{a.w - r.w, a.x - r.x, a.y - r.y, a.z - r.z};
}
__device__ __forceinline__ double partial64(const q64& b) {
#pragma clang fp contract(off)
  return __builtin_fma(b.z, b.z, __builtin_fma(b.y, b.y, b.x * b.x));
}
__device__ __forceinline__ q64 square64(const q64& b, double partial) {
#pragma clang fp contract(off)
  const double t = 2.0 * b.w;
  return q64{__builtin_fma(b.w, b.w, -partial), t * b.x, t * b.y, t * b.z};
}
// operator*(QUATERNION, QUATERNION) (QUATERNION.cpp:154-162), same term order
__device__ __forceinline__ q64 qmul64x(const q64& l, const q64& r) {
#pragma clang fp contract(off)
  q64 f;
  f.x = __builtin_fma(l.w, r.x, __builtin_fma(r.w, l.x, __builtin_fma(-l.z, r.y, l.y * r.z)));
  f.y = __builtin_fma(l.w, r.y, __builtin_fma(r.w, l.y, __builtin_fma(-l.x, r.z, l.z * r.x)));
  f.z = __builtin_fma(l.w, r.z, __builtin_fma(r.w, l.z, __builtin_fma(-l.y, r.x, l.x * r.y)));
  f.w = __builtin_fma(-l.z, r.z, __builtin_fma(-r.y, l.y, __builtin_fma(-l.x, r.x, l.w * r.w)));
  return f;
}

__device__ __forceinline__ bool is_one(double e) { return __builtin_bit_cast(unsigned long long, e) == 0x3FF0000000000000ull; }
__device__ __forceinline__ bool is_two(double e) { return __builtin_bit_cast(unsigned long long, e) == 0x4000000000000000ull; }

// (q - r)^e with the special cases (guarded form).
template <int kPow>
__device__ __forceinline__ q64 term64(const RootRec64& r, const q64& q) {
  if (kPow == kPolar) return qpow(qsub(q, q64{r.w, r.x, r.y, r.z}), r.e);
  const q64 b = qsub64x(q, r);
  if (kPow == kExactOnly || is_one(r.e) || is_two(r.e)) {
    const double partial = partial64(b);
    if (!(partial + b.w * b.w > 0.0)) {
      const double nan = __builtin_nan("");
      return q64{nan, nan, nan, nan};
    }
    if (!(partial > 0.0)) {
      const double m = is_one(r.e) ? fabs(b.w) : b.w * b.w;
      return q64{m, 0.0 * b.x, 0.0 * b.y, 0.0 * b.z};
    }
    if (is_one(r.e)) return b;
    return square64(b, partial);
  }
  return qpow(b, r.e);
}

// (q - r)^e without the special cases; vx keeps the smallest |b.x|. With
// |b.x| >= kExactMinX for every root, partial >= 1e-300 > 0 and term64 takes
// exactly these branches.
constexpr double kExactMinX = 1e-150;
// vx = min(vx, |x|) in one instruction (fmin would canonicalise both inputs
// first); a NaN x leaves vx unchanged, and a NaN product is caught by has_nan.
__device__ __forceinline__ void min_abs64(double& vx, double x) {
  asm("v_min_f64 %0, %1, |%2|" : "=v"(vx) : "v"(vx), "v"(x));
}
template <int kPow>
__device__ __forceinline__ q64 term64_plain(const RootRec64& r, const q64& q, double& vx) {
  const q64 b = qsub64x(q, r);
  min_abs64(vx, b.x);
  if (is_one(r.e)) return b;
  if (kPow == kExactOnly || is_two(r.e)) return square64(b, partial64(b));
  return qpow(b, r.e);
}
// t <- t * (q - r)^e, the product inside each kind's branch (no copies at the merge)
template <int kPow>
__device__ __forceinline__ void apply64_plain(q64& t, const RootRec64& r, const q64& q, double& vx) {
  const q64 b = qsub64x(q, r);
  min_abs64(vx, b.x);
  if (is_one(r.e)) {
    t = qmul64x(t, b);
  } else if (kPow == kExactOnly || is_two(r.e)) {
    t = qmul64x(t, square64(b, partial64(b)));
  } else {
    t = qmul64x(t, qpow(b, r.e));
  }
}

template <int kPow, class RP>
__device__ __forceinline__ q64 poly64_guarded(RP r, int n, const q64& q) {
  q64 res = term64<kPow>(load_root64(r), q);
  for (int i = 1; i < n; i++) {
    res = kPow == kPolar ? qmul(res, term64<kPow>(load_root64(r + i), q))
                         : qmul64x(res, term64<kPow>(load_root64(r + i), q));
    if ((i & 7) == 0 && has_nan(res)) {
      const double nan = __builtin_nan("");
      return q64{nan, nan, nan, nan};
    }
  }
  return res;
}

// The factored product. The exact modes first run it without the special
// cases (no per-term tests, no NaN stand-ins), and only a point where some
// |b.x| < kExactMinX (a special case may apply) runs the guarded form again;
// otherwise the two are the same arithmetic.
template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ q64 poly64(RP r, int n, const q64& q) {
  if (kPow == kPolar) return poly64_guarded<kPow>(r, n, q);
  double vx = __builtin_inf();
  q64 res = term64_plain<kPow>(load_root64(r), q, vx);
  bool nan = false;
  for (int i = 1; i < n; i++) {
    apply64_plain<kPow>(res, load_root64(r + i), q, vx);
    if ((i & 7) == 0 && has_nan(res)) {
      nan = true;
      break;
    }
  }
  if (!(vx >= kExactMinX)) return poly64_guarded<kPow>(r, n, q);
  if (nan) {
    const double n4 = __builtin_nan("");
    return q64{n4, n4, n4, n4};
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
template <int kPow = kPolar, class RP = const RootRec64*, bool kLagrange = true>
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
    // Lagrange order reaches eval64 only through the FP64 method (the API
    // rejects it for the FP32 band, and the mesher evaluates it itself), so
    // the exact-power instances carry no Lagrange code
    if (kPow == kPolar && kLagrange && (a.flags & QJ_MAP_LAGRANGE)) {
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


// ---------------------------------------------------------------- extended range

// A quaternion times 2^e. Normalised so the largest |component| is in [0.5, 1).
struct xq {
  q64 m;
  int e;
};
__device__ __forceinline__ void xnorm(xq& v) {
  const double mx = fmax(fmax(fabs(v.m.w), fabs(v.m.x)), fmax(fabs(v.m.y), fabs(v.m.z)));
  if (mx > 0.0 && isfinite(mx)) {
    int k;
    (void)frexp(mx, &k);
    v.m.w = ldexp(v.m.w, -k);
    v.m.x = ldexp(v.m.x, -k);
    v.m.y = ldexp(v.m.y, -k);
    v.m.z = ldexp(v.m.z, -k);
    v.e += k;
  }
}
// QUATERNION::pow (QUATERNION.cpp:375-393) with exps = exp(e log|q| + lnoff)
// kept as mantissa * 2^k; lnoff carries the scale of a normalised base
// (e k ln 2 for a base m 2^k: the angle does not depend on the scale).
__device__ __forceinline__ xq xpow(const q64& q, double e, double lnoff = 0.0) {
  const double partial = q.x * q.x + q.y * q.y + q.z * q.z;
  const double qMagnitude = sqrt(partial + q.w * q.w);
  const double vMagnitude = sqrt(partial);
  const double vMagnitudeInv = (vMagnitude > 0.0) ? 1.0 / vMagnitude : 0.0;
  const double scale = e * acos(q.w / qMagnitude) * vMagnitudeInv;
  const double magnitude = scale * vMagnitude;
  const double magnitudeInv = (magnitude > 0.0) ? 1.0 / magnitude : 0.0;
  const double lnE = e * log(qMagnitude) + lnoff;
  int k = 0;
  double em;
  if (isfinite(lnE)) {
    k = (int)floor(lnE * 1.4426950408889634);
    em = exp(lnE - k * 0.6931471805599453);
  } else {
    em = exp(lnE);
  }
  double s, c;
  sincos(magnitude, &s, &c);
  const double scale2 = scale * em * magnitudeInv * s;
  xq r{{em * c, scale2 * q.x, scale2 * q.y, scale2 * q.z}, k};
  xnorm(r);
  return r;
}

// (q - r)^e with a separate binary exponent, for QJ_MAP_EXTENDED_RANGE: the
// base is normalised first, so neither its square nor the polar magnitude
// leaves FP64's range; exponents 1 and 2 (kPow != kPolar) as the exact
// products with the polar form's special cases (b = 0: NaN; v = 0:
// (|w|^e, +0 v)), decided on the normalised base as long double would.
template <int kPow>
__device__ __forceinline__ xq term_ext(const RootRec64& r, const q64& q) {
  xq n{qsub64x(q, r), 0};
  xnorm(n);
  if (kPow == kPolar || !(kPow == kExactOnly || is_one(r.e) || is_two(r.e)))
    return xpow(n.m, r.e, r.e * (double)n.e * 0.6931471805599453);
  const double partial = partial64(n.m);
  if (!(partial + n.m.w * n.m.w > 0.0)) {
    const double nan = __builtin_nan("");
    return xq{{nan, nan, nan, nan}, 0};
  }
  if (!(partial > 0.0)) {
    xq v = is_one(r.e) ? xq{{fabs(n.m.w), 0.0 * n.m.x, 0.0 * n.m.y, 0.0 * n.m.z}, n.e}
                       : xq{{n.m.w * n.m.w, 0.0 * n.m.x, 0.0 * n.m.y, 0.0 * n.m.z}, 2 * n.e};
    xnorm(v);
    return v;
  }
  if (is_one(r.e)) return n;
  xq v{square64(n.m, partial), 2 * n.e};
  xnorm(v);
  return v;
}

template <int kPow, class RP>
__device__ __forceinline__ xq poly_ext(RP r, int n, const q64& q) {
  xq res = term_ext<kPow>(load_root64(r), q);
  for (int i = 1; i < n; i++) {
    const xq t = term_ext<kPow>(load_root64(r + i), q);
    res.m = qmul64x(res.m, t.m);
    res.e += t.e;
    xnorm(res);
  }
  return res;
}

// One voxel with the range of the reference's default build (Real = long
// double, SETTINGS.h:10; QJ_MAP_EXTENDED_RANGE): the factored products, the
// quotient and |q| are FP64 mantissas with separate binary exponents, so
// nothing overflows or underflows where x87 long double does not. The
// escape and black-hole tests compare the FP64 value of |q| (exact whenever
// it is representable; +inf or 0 beyond, which decides them the same way);
// F = log|mantissa| + exponent ln 2. An iterate that continues below FP64's
// range (no black-hole stop) continues from its FP64 value.
template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ Res64 eval64e(const MapArgs& a, RP roots, q64 q) {
  Res64 r{0.0, 0, true, false};
  double magnitude = qmag(q);
  if ((a.flags & QJ_MAP_SKIP_ORIGIN) && magnitude < 1e-8) {
    r.written = false;
    return r;
  }
  double logmag = log(magnitude);
  int it = 0;
  while (magnitude < a.escape && it < a.max_iter) {
    xq v = poly_ext<kPow>(roots, a.nt, q);
    if (a.nb > 0) {
      const xq B = poly_ext<kPow>(roots + a.nt, a.nb, q);
      v.m = qdiv(v.m, B.m);
      v.e -= B.e;
    }
    v.m = qscale(v.m, a.exp_scaling);
    xnorm(v);
    const double mm = qmag(v.m);
    logmag = log(mm) + (double)v.e * 0.6931471805599453;
    magnitude = ldexp(mm, v.e);
    q = q64{ldexp(v.m.w, v.e), ldexp(v.m.x, v.e), ldexp(v.m.y, v.e), ldexp(v.m.z, v.e)};
    it++;
    r.iters++;
    if ((a.flags & QJ_MAP_BLACK_HOLE) && magnitude < kBlackHole) it = a.max_iter;
  }
  r.F = logmag;
  return r;
}

// eval64 or, with QJ_MAP_EXTENDED_RANGE, eval64e (a uniform branch where a
// wave's voxels share their map arguments)
template <int kPow = kPolar, class RP = const RootRec64*>
__device__ __forceinline__ Res64 eval64_any(const MapArgs& a, RP roots, const q64& q) {
  if (a.flags & QJ_MAP_EXTENDED_RANGE) return eval64e<kPow>(a, roots, q);
  return eval64<kPow>(a, roots, q);
}

}  // namespace qj
