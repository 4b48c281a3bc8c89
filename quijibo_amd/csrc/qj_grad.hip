// Gradient kernels.
//
// k_ap_grads: the one-iteration AP score and exponent gradients of the
//   optimizer (OPTIMIZE_3D.cpp computeScoreAP :3563, computeExpGradientAP
//   :3589-3625, computeTopGradientAP :8905-9106, computeBottomGradientAP
//   :10012, computeTop/BottomBulkScalarGradientAP :9250/:10360). One lane per
//   cell; every root's term is reduced over the block (wave shuffles + LDS)
//   into per-block partials, summed in a fixed order by k_sum_partials, so the
//   result is deterministic. The reference keeps per-cell prefix/suffix caches
//   (POLYNOMIAL_CACHE); here nothing is stored per cell: a forward sweep forms
//   T and B, and a backward sweep re-forms each term, carries the suffix
//   product S and recovers the prefix as T * S^-1 (quaternion products are
//   conformal, so this costs no accuracy).
//   Cells that overflow FP64 are redone with a shared binary exponent per
//   quaternion (the reference redoes them in long double, :9011-9103).
//
// k_adjoint: reverse mode through max_iter iterations of the map at points,
//   recomputing every iteration's products instead of storing trajectories.
//   Each point is a group of 8 lanes, each lane taking a share of the roots;
//   the ordered product of the shares is a prefix/suffix scan of quaternion
//   products over the group, and each root's gradient is summed over the
//   wave's points by shuffles before one lane adds it to the block's LDS.
#include <float.h>

#include <string.h>

#include <type_traits>

#include "qj_geom.h"
#include "qj_internal.h"
#include "qj_quat.h"

namespace qj {

constexpr int kGBlock = 256;

// ------------------------------------------------------------ extended range

struct xq {
  q64 m;
  int e;
};
__device__ __forceinline__ xq xnorm(xq v) {
  const double mx = fmax(fmax(fabs(v.m.w), fabs(v.m.x)), fmax(fabs(v.m.y), fabs(v.m.z)));
  if (mx > 0.0 && isfinite(mx)) {
    int k;
    (void)frexp(mx, &k);
    v.m = q64{ldexp(v.m.w, -k), ldexp(v.m.x, -k), ldexp(v.m.y, -k), ldexp(v.m.z, -k)};
    v.e += k;
  }
  return v;
}
__device__ __forceinline__ xq xfrom(const q64& q) { return xnorm(xq{q, 0}); }
__device__ __forceinline__ xq mul(const xq& a, const xq& b) { return xnorm(xq{qmul(a.m, b.m), a.e + b.e}); }
__device__ __forceinline__ xq inv(const xq& a) {
  const q64 i = qinv(a.m);
  return xnorm(xq{i, -a.e});
}
__device__ __forceinline__ xq scaled(const xq& a, double s) { return xnorm(xq{qscale(a.m, s), a.e}); }
__device__ __forceinline__ double logmag(const xq& a) { return log(qmag(a.m)) + a.e * 0.6931471805599453; }
// <a,b> / <c,c> as a double
__device__ __forceinline__ double dotratio(const xq& a, const xq& b, const xq& c) {
  return qdot(a.m, b.m) / qdot(c.m, c.m) * ldexp(1.0, a.e + b.e - 2 * c.e);
}
__device__ __forceinline__ bool ok(const xq& a) {
  return isfinite(a.m.w) && isfinite(a.m.x) && isfinite(a.m.y) && isfinite(a.m.z);
}
// QUATERNION::pow with exps = exp(e log|q|) kept as mantissa * 2^k
__device__ __forceinline__ xq pw_x(const q64& q, double e) {
  const double partial = q.x * q.x + q.y * q.y + q.z * q.z;
  const double qMagnitude = sqrt(partial + q.w * q.w);
  const double vMagnitude = sqrt(partial);
  const double vMagnitudeInv = (vMagnitude > 0.0) ? 1.0 / vMagnitude : 0.0;
  const double scale = e * acos(q.w / qMagnitude) * vMagnitudeInv;
  const double magnitude = scale * vMagnitude;
  const double magnitudeInv = (magnitude > 0.0) ? 1.0 / magnitude : 0.0;
  const double lnE = e * log(qMagnitude);
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
  return xnorm(xq{{em * c, scale2 * q.x, scale2 * q.y, scale2 * q.z}, k});
}

// plain FP64 counterparts with the same interface
__device__ __forceinline__ q64 mul(const q64& a, const q64& b) { return qmul(a, b); }
__device__ __forceinline__ q64 inv(const q64& a) { return qinv(a); }
__device__ __forceinline__ q64 scaled(const q64& a, double s) { return qscale(a, s); }
__device__ __forceinline__ double logmag(const q64& a) { return log(qmag(a)); }
__device__ __forceinline__ double dotratio(const q64& a, const q64& b, const q64& c) {
  return (1.0 / qdot(c, c)) * qdot(a, b);  // (1.0 / R.dot(R)) * R.dot(Rd), OPTIMIZE_3D.cpp:8976
}
__device__ __forceinline__ bool ok(const q64& a) {
  return isfinite(a.w) && isfinite(a.x) && isfinite(a.y) && isfinite(a.z);
}
template <class Q>
__device__ __forceinline__ Q from64(const q64& q);
template <>
__device__ __forceinline__ q64 from64<q64>(const q64& q) { return q; }
template <>
__device__ __forceinline__ xq from64<xq>(const q64& q) { return xfrom(q); }
template <class Q>
__device__ __forceinline__ Q pw(const q64& b, double e);
template <>
__device__ __forceinline__ q64 pw<q64>(const q64& b, double e) { return qpow(b, e); }
template <>
__device__ __forceinline__ xq pw<xq>(const q64& b, double e) { return pw_x(b, e); }

// ------------------------------------------------------------ block reduction

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// acc[wave * K + k] += wave sum of v (lane 0 writes); K slots per wave
__device__ __forceinline__ void wave_acc(double* acc, int K, int k, double v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) acc[(threadIdx.x >> 6) * K + k] += v;
}

// ------------------------------------------------------------ AP gradients

struct APArgs {
  MapArgs a;
  GridGeom wg;
  double alpha, thr;
};

// slot layout of the partial sums
enum { SL_SCORE = 0, SL_GEXP = 1, SL_GTS = 2, SL_GBS = 3, SL_ROOTS = 4 };

// A term that came out non-finite in FP64 redone with explicit exponents from
// the same FP64 factors (the reference redoes it in long double,
// OPTIMIZE_3D.cpp:8986-8995, 9011-9103). Returns 0 if a factor itself is not finite.
__device__ __forceinline__ double top_term_x(const q64& left, const q64& power, double scalar, const q64& lg,
                                             const q64& right, const q64& T, const q64& B, double dSdF) {
  if (!(ok(left) && ok(power) && ok(right) && ok(T) && ok(B))) return 0.0;
  const xq R = mul(xfrom(T), inv(xfrom(B)));
  const xq mid = mul(scaled(xfrom(power), scalar), xfrom(lg));
  const xq Rd = mul(mul(mul(xfrom(left), mid), xfrom(right)), inv(xfrom(B)));
  const double v = dSdF * dotratio(R, Rd, R);
  return isfinite(v) ? v : 0.0;
}
__device__ __forceinline__ double bottom_term_x(const q64& left, const q64& power, double scalar, const q64& ilg,
                                                const q64& right, const q64& T, const q64& B, double dSdF) {
  if (!(ok(left) && ok(power) && ok(right) && ok(T) && ok(B) && ok(ilg))) return 0.0;
  const xq R = mul(xfrom(T), inv(xfrom(B)));
  const xq mid = mul(scaled(xfrom(power), scalar), xfrom(ilg));
  const xq bd = mul(mul(mul(xfrom(q64{-1, 0, 0, 0}), xfrom(left)), mid), xfrom(right));
  const double v = dSdF * dotratio(R, mul(xfrom(T), inv(bd)), R);
  return isfinite(v) ? v : 0.0;
}

// Accumulates one cell's root terms (top: per root and bulk; bottom: per root
// and bulk) into acc. Every lane of the wave calls this (the reductions are
// wave-wide); lanes with active == false add zeros.
template <class Q>
__device__ __forceinline__ void cell_terms(const APArgs& g, const RootRec64* __restrict__ roots, const q64& q0,
                                           const Q& T, const Q& B, double w, bool active, double* acc, int K) {
  const MapArgs& a = g.a;
  Q R = T;
  double dSdF = 0.0;
  bool valid = false;
  if (active) {
    R = mul(T, inv(B));
    const double Rm = logmag(R);
    // |R| <= 0 or inf: the reference skips the cell (OPTIMIZE_3D.cpp:8956)
    valid = isfinite(Rm) && ok(R);
    const double F = Rm + log(a.exp_scaling);
    const double tanhTerm = tanh(g.alpha * (F - g.thr));
    dSdF = -g.alpha * w * (1 - tanhTerm * tanhTerm);
  }
  const Q one = from64<Q>(q64{1, 0, 0, 0});
  // top: backward sweep, S = suffix product, prefix = T * S^-1
  {
    Q S = one;
    for (int i = a.nt - 1; i >= 0; i--) {
      double vt = 0.0, vb = 0.0;
      if (valid) {
        const RootRec64 r = roots[i];
        const q64 base = qsub(q0, q64{r.w, r.x, r.y, r.z});
        const Q power = pw<Q>(base, r.e);
        const Q right = S;  // S_{i+1}
        S = mul(power, S);  // S_i
        const Q left = (i > 0) ? mul(T, inv(S)) : one;
        const q64 lg = qlog(base);
        const Q iB = inv(B);
        vt = dSdF * dotratio(R, mul(mul(mul(left, mul(scaled(power, a.sc_top), from64<Q>(lg))), right), iB), R);
        vb = dSdF * dotratio(R, mul(mul(mul(left, mul(scaled(power, r.p), from64<Q>(lg))), right), iB), R);
        if constexpr (std::is_same<Q, q64>::value) {
          if (!isfinite(vt)) vt = top_term_x(left, power, a.sc_top, lg, right, T, B, dSdF);
          if (!isfinite(vb)) vb = top_term_x(left, power, r.p, lg, right, T, B, dSdF);
        } else {
          if (!isfinite(vt)) vt = 0.0;
          if (!isfinite(vb)) vb = 0.0;
        }
      }
      wave_acc(acc, K, SL_ROOTS + i, vt);
      wave_acc(acc, K, SL_GTS, vb);
    }
  }
  if (a.nb > 0) {
    const RootRec64* rb = roots + a.nt;
    Q S = one;
    for (int i = a.nb - 1; i >= 0; i--) {
      double vt = 0.0, vb = 0.0;
      if (valid) {
        const RootRec64 r = rb[i];
        const q64 base = qsub(q0, q64{r.w, r.x, r.y, r.z});
        const Q power = pw<Q>(base, r.e);
        const Q right = S;
        S = mul(power, S);
        const Q left = (i > 0) ? mul(B, inv(S)) : one;
        const q64 ilg = qinv(qlog(base));
        const Q m1 = from64<Q>(q64{-1, 0, 0, 0});
        if (a.sc_bottom > 0.0) {  // OPTIMIZE_3D.cpp:10084
          const Q bd = mul(mul(mul(m1, left), mul(scaled(power, 1.0 / a.sc_bottom), from64<Q>(ilg))), right);
          vt = dSdF * dotratio(R, mul(T, inv(bd)), R);
          if constexpr (std::is_same<Q, q64>::value) {
            if (!isfinite(vt)) vt = bottom_term_x(left, power, 1.0 / a.sc_bottom, ilg, right, T, B, dSdF);
          } else if (!isfinite(vt)) {
            vt = 0.0;
          }
        }
        if (!(r.p < 1e-8)) {  // OPTIMIZE_3D.cpp:10404-10405
          const Q bd = mul(mul(mul(m1, left), mul(scaled(power, 1.0 / r.p), from64<Q>(ilg))), right);
          vb = dSdF * dotratio(R, mul(T, inv(bd)), R);
          if constexpr (std::is_same<Q, q64>::value) {
            if (!isfinite(vb)) vb = bottom_term_x(left, power, 1.0 / r.p, ilg, right, T, B, dSdF);
          } else if (!isfinite(vb)) {
            vb = 0.0;
          }
        }
      }
      wave_acc(acc, K, SL_ROOTS + a.nt + i, vt);
      wave_acc(acc, K, SL_GBS, vb);
    }
  }
}

template <class Q>
__device__ __forceinline__ Q poly_fwd(const RootRec64* __restrict__ r, int n, const q64& q) {
  Q res = pw<Q>(qsub(q, q64{r[0].w, r[0].x, r[0].y, r[0].z}), r[0].e);
  for (int i = 1; i < n; i++) res = mul(res, pw<Q>(qsub(q, q64{r[i].w, r[i].x, r[i].y, r[i].z}), r[i].e));
  return res;
}

__global__ __launch_bounds__(kGBlock) void k_ap_grads(APArgs g, const RootRec64* __restrict__ roots,
                                                      const double* __restrict__ weight, double* __restrict__ field,
                                                      double* __restrict__ partials, int K,
                                                      unsigned long long* __restrict__ n_adaptive) {
  extern __shared__ double acc[];  // (kGBlock/64) x K
  const MapArgs& a = g.a;
  for (int i = threadIdx.x; i < (kGBlock / 64) * K; i += kGBlock) acc[i] = 0.0;
  __syncthreads();
  const int64_t X = a.res[0], XY = X * a.res[1], N = XY * a.res[2];
  const int64_t idx = (int64_t)blockIdx.x * kGBlock + threadIdx.x;
  bool active = idx < N;
  double c[3] = {0, 0, 0};
  double w = 0.0;
  q64 q0{0, 0, 0, 0};
  double Ffield = 0.0;
  bool iterate = false;
  if (active) {
    const int z = (int)(idx / XY), y = (int)((idx - z * XY) / X), x = (int)(idx - z * XY - (int64_t)y * X);
    cell_center(a, x, y, z, c);
    w = field_lookup(g.wg, weight, c);
    q0 = q64{c[0], c[1], c[2], a.slice};
    const double mag0 = qmag(q0);
    if (mag0 < 1e-8) {
      Ffield = field[idx];  // unwritten cell: score terms read the existing value
    } else if (!(mag0 < a.escape) || a.max_iter < 1) {
      Ffield = log(mag0);
      field[idx] = Ffield;
    } else {
      iterate = true;
    }
  }
  // forward sweep (computeRationalMapAP with one iteration)
  q64 T{1, 0, 0, 0}, B{1, 0, 0, 0};
  bool ovf = false;
  if (iterate) {
    T = poly_fwd<q64>(roots, a.nt, q0);
    q64 qn;
    if (a.nb > 0) {
      B = poly_fwd<q64>(roots + a.nt, a.nb, q0);
      qn = qdiv(T, B);
      const double m2 = B.w * B.w + B.x * B.x + B.y * B.y + B.z * B.z;
      ovf = !(ok(T) && ok(B) && isfinite(m2) && ok(qn));
    } else {
      qn = T;
      ovf = !ok(T);
    }
    qn = qscale(qn, a.exp_scaling);
    const double mag = qmag(qn);
    ovf |= !isfinite(mag);
    Ffield = log(mag);
  }
  // gradients in FP64 for the cells that stayed finite
  cell_terms<q64>(g, roots, q0, T, B, w, iterate && !ovf, acc, K);
  // overflowed cells: the whole cell with explicit binary exponents
  const bool redo = iterate && ovf;
  if (__ballot(redo)) {
    xq TX = xfrom(q64{1, 0, 0, 0}), BX = TX;
    if (redo) {
      TX = poly_fwd<xq>(roots, a.nt, q0);
      if (a.nb > 0) BX = poly_fwd<xq>(roots + a.nt, a.nb, q0);
      const xq R = scaled(mul(TX, inv(BX)), a.exp_scaling);
      const double l = logmag(R);
      Ffield = isnan(l) ? 0.0 : l;  // OPTIMIZE_3D.cpp:1557-1561
    }
    cell_terms<xq>(g, roots, q0, TX, BX, w, redo, acc, K);
  }
  if (iterate) field[idx] = Ffield;
  double sc = 0.0, ge = 0.0;
  if (active) {
    sc = -tanh(g.alpha * Ffield) * w;
    const double S = tanh(g.alpha * (Ffield - g.thr));
    ge = -g.alpha * w * (1.0 - S * S);
  }
  wave_acc(acc, K, SL_SCORE, sc);
  wave_acc(acc, K, SL_GEXP, ge);
  const unsigned long long nr = __popcll(__ballot(redo));
  if ((threadIdx.x & 63) == 0 && nr) atomicAdd(n_adaptive, nr);
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += kGBlock) {
    double s = 0.0;
    for (int wv = 0; wv < kGBlock / 64; wv++) s += acc[wv * K + k];
    partials[(int64_t)blockIdx.x * K + k] = s;
  }
}

// sums partials[b][k] over b: one block per k, each thread a strided run of
// blocks in order, then a fixed tree over the threads (deterministic; a serial
// sum over thousands of blocks per k took most of a millisecond)
constexpr int kSumThreads = 256;
__global__ __launch_bounds__(kSumThreads) void k_sum_partials(const double* __restrict__ partials, int64_t nb, int K,
                                                              double* __restrict__ out) {
  __shared__ double s_v[kSumThreads];
  const int k = blockIdx.x;
  double v = 0.0;
  for (int64_t b = threadIdx.x; b < nb; b += kSumThreads) v += partials[b * K + k];
  s_v[threadIdx.x] = v;
  __syncthreads();
  for (int w = kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s_v[threadIdx.x] += s_v[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[k] = s_v[0];
}

qj_status ap_grads(qj_workspace* ws, const MapArgs& a, const qj_rational* f, const qj_grid* wg, const double* d_weight,
                   double alpha, double thr, double* d_field, double* gt, double* gb, qj_ap_grads* out,
                   hipStream_t s) {
  const int64_t N = (int64_t)a.res[0] * a.res[1] * a.res[2];
  const int K = SL_ROOTS + a.nt + a.nb;
  const int64_t nblocks = (N + kGBlock - 1) / kGBlock;
  const size_t lds = sizeof(double) * (kGBlock / 64) * K;
  if (lds > 64 * 1024) {
    set_error("too many roots for the gradient kernel (%d)", a.nt + a.nb);
    return QJ_ERR_CAPACITY;
  }
  qj_status st = ensure_partials(ws, nblocks * K + K + 1, s);
  if (st) return st;
  APArgs g;
  g.a = a;
  make_geom(wg, &g.wg);
  g.alpha = alpha;
  g.thr = thr;
  double* sums = ws->d_partials + nblocks * K;
  unsigned long long* nad = (unsigned long long*)(sums + K);
  QJ_HIP(hipMemsetAsync(nad, 0, sizeof(unsigned long long), s));
  hipLaunchKernelGGL(k_ap_grads, dim3((unsigned)nblocks), dim3(kGBlock), lds, s, g, f->d64, d_weight, d_field,
                     ws->d_partials, K, nad);
  QJ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_sum_partials, dim3(K), dim3(kSumThreads), 0, s, ws->d_partials, nblocks, K, sums);
  QJ_HIP(hipGetLastError());
  std::vector<double> h(K + 1);
  QJ_HIP(hipMemcpyAsync(h.data(), sums, sizeof(double) * (K + 1), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipStreamSynchronize(s));
  out->score = h[SL_SCORE];
  out->grad_exp = h[SL_GEXP];
  out->grad_top_scalar = h[SL_GTS];
  out->grad_bottom_scalar = h[SL_GBS];
  unsigned long long na;
  ::memcpy(&na, &h[K], sizeof(na));
  out->adaptive_cells = (int64_t)na;
  for (int i = 0; i < a.nt; i++) gt[i] = h[SL_ROOTS + i];
  for (int i = 0; i < a.nb; i++) gb[i] = h[SL_ROOTS + a.nt + i];
  return QJ_OK;
}

// ------------------------------------------------------------ adjoint at points

template <class T>
struct Q4 {
  T w, x, y, z;
};
template <class T>
__device__ __forceinline__ Q4<T> mul(const Q4<T>& l, const Q4<T>& r) {
  Q4<T> f;
  f.x = l.y * r.z - l.z * r.y + r.w * l.x + l.w * r.x;
  f.y = l.z * r.x - l.x * r.z + r.w * l.y + l.w * r.y;
  f.z = l.x * r.y - l.y * r.x + r.w * l.z + l.w * r.z;
  f.w = l.w * r.w - l.x * r.x - r.y * l.y - l.z * r.z;
  return f;
}
template <class T>
__device__ __forceinline__ Q4<T> conj(const Q4<T>& a) { return {a.w, -a.x, -a.y, -a.z}; }
template <class T>
__device__ __forceinline__ Q4<T> sc(const Q4<T>& a, T s) { return {a.w * s, a.x * s, a.y * s, a.z * s}; }
template <class T>
__device__ __forceinline__ T dot(const Q4<T>& a, const Q4<T>& b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T>
__device__ __forceinline__ Q4<T> unit(const Q4<T>& a) {
  const T n = sqrt(dot(a, a));
  return n > T(0) ? sc(a, T(1) / n) : a;
}

// One polar power term in unit form: t = |b|^e * that, with that = (cos m, sin m v/|v|),
// m = e * atan2(|v|, w) (QUATERNION::pow, QUATERNION.cpp:375-393, for e > 0).
__device__ __forceinline__ void sincos_t(double m, double* s, double* c) { sincos(m, s, c); }
__device__ __forceinline__ void sincos_t(float m, float* s, float* c) { sincosf(m, s, c); }

template <class T>
struct Term {
  Q4<T> b;
  Q4<T> that;
  T r, rho, th, lb;  // |b|, |v|, angle, log|b|
  T A, C;            // that = (A, C v)
};
template <class T>
__device__ __forceinline__ Term<T> make_term(const Q4<T>& q, const RootRec64& rr, T e) {
  Term<T> t;
  t.b = Q4<T>{q.w - T(rr.w), q.x - T(rr.x), q.y - T(rr.y), q.z - T(rr.z)};
  const T p = t.b.x * t.b.x + t.b.y * t.b.y + t.b.z * t.b.z;
  t.rho = sqrt(p);
  t.r = sqrt(p + t.b.w * t.b.w);
  t.th = atan2(t.rho, t.b.w);
  t.lb = log(t.r);
  const T m = e * t.th;
  T s, c;
  sincos_t(m, &s, &c);
  t.A = c;
  t.C = t.rho > T(0) ? s / t.rho : e / t.b.w;  // limit v -> 0 (w > 0)
  t.that = Q4<T>{t.A, t.C * t.b.x, t.C * t.b.y, t.C * t.b.z};
  return t;
}
// b_bar = J'^T tbar' for the unit-form power (J = |b|^e J')
template <class T>
__device__ __forceinline__ Q4<T> pow_vjp(const Term<T>& t, const Q4<T>& tb, T e) {
  const T r = t.r, rho = t.rho, w = t.b.w;
  const T gA = tb.w;
  const T gC = tb.x * t.b.x + tb.y * t.b.y + tb.z * t.b.z;
  const T Gr = (gA * e * t.A + gC * e * t.C) / r;
  if (!(rho > T(0))) {
    // purely real base: d/db of (w^e, e w^(e-1) v) / w^e
    return Q4<T>{Gr, t.C * tb.x, t.C * tb.y, t.C * tb.z};
  }
  const T Gth = -gA * e * t.C * rho + gC * e * t.A / rho;
  const T Grho = -gC * t.C / rho;
  const T r2 = r * r;
  const T coef = Gr / r + Gth * w / (r2 * rho) + Grho / rho;
  return Q4<T>{Gr * (w / r) - Gth * (rho / r2), coef * t.b.x + t.C * tb.x, coef * t.b.y + t.C * tb.y,
               coef * t.b.z + t.C * tb.z};
}

constexpr int kAdjBlock = 128;
constexpr int kMaxAdjIter = 16;
// lanes per point: each lane of a group takes a contiguous share of the top
// roots and of the bottom roots, so 40k points keep every SIMD busy
template <class T>
constexpr int adj_lanes() {
  return sizeof(T) == 4 ? 16 : 8;  // FP64: 8 (twice the registers per value, 16 measured slower)
}

template <class T>
__device__ __forceinline__ Q4<T> shfl_up_q(const Q4<T>& a, int d) {
  return {__shfl_up(a.w, d, adj_lanes<T>()), __shfl_up(a.x, d, adj_lanes<T>()), __shfl_up(a.y, d, adj_lanes<T>()),
          __shfl_up(a.z, d, adj_lanes<T>())};
}
template <class T>
__device__ __forceinline__ Q4<T> shfl_down_q(const Q4<T>& a, int d) {
  return {__shfl_down(a.w, d, adj_lanes<T>()), __shfl_down(a.x, d, adj_lanes<T>()), __shfl_down(a.y, d, adj_lanes<T>()),
          __shfl_down(a.z, d, adj_lanes<T>())};
}
template <class T>
__device__ __forceinline__ Q4<T> shfl_q(const Q4<T>& a, int src) {
  return {__shfl(a.w, src, adj_lanes<T>()), __shfl(a.x, src, adj_lanes<T>()), __shfl(a.y, src, adj_lanes<T>()),
          __shfl(a.z, src, adj_lanes<T>())};
}
// sum over the lanes of a group
template <class T>
__device__ __forceinline__ T group_sum(T v) {
  for (int o = adj_lanes<T>() / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
template <class T>
__device__ __forceinline__ Q4<T> group_sum_q(const Q4<T>& a) {
  return {group_sum(a.w), group_sum(a.x), group_sum(a.y), group_sum(a.z)};
}
// sum over the groups of a wave, lane by lane (the lanes that hold the same roots)
template <class T>
__device__ __forceinline__ double across_groups(double v) {
  for (int o = adj_lanes<T>(); o < 64; o <<= 1) v += __shfl_xor(v, o);
  return v;
}

// The unit product of polynomial roots [lo, hi) in order, its log magnitude,
// and over the group: the whole polynomial's unit product and log magnitude,
// and the unit product of the roots after this lane's share.
template <class T>
struct Share {
  Q4<T> all, after;
  T L;
};
template <class T>
__device__ __forceinline__ Share<T> group_product(const RootRec64* __restrict__ roots, int lo, int hi, const Q4<T>& q,
                                                  int sub) {
  Q4<T> p{1, 0, 0, 0};
  T l = 0;
  for (int i = lo; i < hi; i++) {
    const RootRec64 rr = roots[i];
    const Term<T> t = make_term(q, rr, T(rr.e));
    p = mul(p, t.that);
    l += T(rr.e) * t.lb;
    if (((i - lo) & 7) == 7) p = unit(p);
  }
  p = unit(p);
  // ordered inclusive prefix and suffix products over the group (Hillis-Steele)
  Q4<T> pre = p, suf = p;
  for (int d = 1; d < adj_lanes<T>(); d <<= 1) {
    const Q4<T> y = shfl_up_q(pre, d);
    const Q4<T> z = shfl_down_q(suf, d);
    if (sub >= d) pre = mul(y, pre);
    if (sub + d < adj_lanes<T>()) suf = mul(suf, z);
  }
  Share<T> s;
  s.all = unit(shfl_q(pre, adj_lanes<T>() - 1));
  const Q4<T> nxt = shfl_down_q(suf, 1);
  s.after = sub + 1 < adj_lanes<T>() ? unit(nxt) : Q4<T>{1, 0, 0, 0};
  s.L = group_sum(l);
  return s;
}

// slot layout per root i (top then bottom): 5 i + {0..3} = d/d root component, 5 i + 4 = d/d e_i
// One group of adj_lanes<T>() lanes per point; lane `sub` of a group takes top
// roots [sub ct, sub ct + ct) and bottom roots [sub cb, sub cb + cb). Control
// flow is uniform over the wave (every group runs max_iter steps and ct / cb
// root steps, masked), so the shuffles always see active partners.
template <class T>
__global__ __launch_bounds__(kAdjBlock) void k_adjoint(MapArgs a, const RootRec64* __restrict__ roots,
                                                       const double* __restrict__ pts, int64_t n,
                                                       const double* __restrict__ dJdF,
                                                       double* __restrict__ partials, int K) {
  extern __shared__ double acc[];  // (kAdjBlock/64) x K
  for (int i = threadIdx.x; i < (kAdjBlock / 64) * K; i += kAdjBlock) acc[i] = 0.0;
  __syncthreads();
  constexpr int kL = adj_lanes<T>();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int sub = lane & (kL - 1);
  const int64_t pid = ((int64_t)blockIdx.x * kAdjBlock + threadIdx.x) / kL;
  const bool active = pid < n;
  const int ct = (a.nt + kL - 1) / kL, cb = (a.nb + kL - 1) / kL;
  const int t0 = sub * ct, t1 = min(a.nt, t0 + ct);
  const int b0 = sub * cb, b1 = min(a.nb, b0 + cb);
  Q4<T> qs[kMaxAdjIter + 1];
  int Kx = 0;
  T gF = 0;
  qs[0] = active ? Q4<T>{T(pts[3 * pid]), T(pts[3 * pid + 1]), T(pts[3 * pid + 2]), T(a.slice)}
                 : Q4<T>{1, 0, 0, 0};
  if (active) gF = T(dJdF[pid]);
  bool running = active && sqrt(dot(qs[0], qs[0])) < T(a.escape) && a.max_iter > 0;
  for (int k = 0; k < a.max_iter; k++) {
    const Q4<T> q = running ? qs[k] : Q4<T>{1, 0, 0, 0};
    const Share<T> tp = group_product(roots, t0, t1, q, sub);
    const Share<T> bt = group_product(roots + a.nt, b0, b1, q, sub);
    if (running) {
      const Q4<T> Qh = a.nb > 0 ? mul(tp.all, conj(bt.all)) : tp.all;
      const T scale = T(a.exp_scaling) * exp(tp.L - (a.nb > 0 ? bt.L : T(0)));
      qs[k + 1] = sc(Qh, scale);
      Kx = k + 1;
      const T mag = sqrt(dot(qs[k + 1], qs[k + 1]));
      if (!(mag < T(a.escape)) || ((a.flags & QJ_MAP_BLACK_HOLE) && mag < T(10.0 * (double)FLT_MIN)))
        running = false;
    }
  }
  // dJ/dq_K for F = log|q_K|
  Q4<T> qbar{0, 0, 0, 0};
  if (active && gF != T(0)) {
    const Q4<T> qk = qs[Kx];
    qbar = sc(qk, gF / dot(qk, qk));
  }
  for (int k = a.max_iter - 1; k >= 0; k--) {
    const bool live = active && k < Kx;
    const Q4<T> q = live ? qs[k] : Q4<T>{1, 0, 0, 0};
    const Share<T> tp = group_product(roots, t0, t1, q, sub);
    const Share<T> bt = group_product(roots + a.nt, b0, b1, q, sub);
    const Q4<T> Th = tp.all;
    const Q4<T> Bh = a.nb > 0 ? bt.all : Q4<T>{1, 0, 0, 0};
    const T Qmag = exp(tp.L - (a.nb > 0 ? bt.L : T(0)));
    const Q4<T> Qbar = sc(qbar, T(a.exp_scaling));  // q_{k+1} = c Q
    Q4<T> qnew{0, 0, 0, 0};
    // top terms: tbar' = conj(Phat) Qbar Bhat conj(Shat_{i+1})
    {
      const Q4<T> QB = mul(Qbar, Bh);
      Q4<T> S = tp.after;
      for (int j = 0; j < ct; j++) {
        const int i = t0 + ct - 1 - j;
        double g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
        if (live && i < t1) {
          const RootRec64 rr = roots[i];
          const Term<T> t = make_term(q, rr, T(rr.e));
          const Q4<T> right = S;
          S = mul(t.that, S);
          const Q4<T> P = (i > 0) ? mul(Th, conj(S)) : Q4<T>{1, 0, 0, 0};
          const Q4<T> tb = mul(mul(conj(P), QB), conj(right));
          const Q4<T> bb = sc(pow_vjp(t, tb, T(rr.e)), Qmag);
          // d t / d e = t log(b); log(b) = (log|b|, th v/|v|)
          const T lv = t.rho > T(0) ? t.th / t.rho : T(0);
          const Q4<T> lg{t.lb, lv * t.b.x, lv * t.b.y, lv * t.b.z};
          const T ge = Qmag * dot(tb, mul(t.that, lg));
          qnew = Q4<T>{qnew.w + bb.w, qnew.x + bb.x, qnew.y + bb.y, qnew.z + bb.z};
          g0 = -(double)bb.w; g1 = -(double)bb.x; g2 = -(double)bb.y; g3 = -(double)bb.z; g4 = (double)ge;
          if ((j & 7) == 7) S = unit(S);
        }
        g0 = across_groups<T>(g0);
        g1 = across_groups<T>(g1);
        g2 = across_groups<T>(g2);
        g3 = across_groups<T>(g3);
        g4 = across_groups<T>(g4);
        if (lane < kL && i < t1) {
          double* r = acc + wave * K + 5 * i;
          r[0] += g0; r[1] += g1; r[2] += g2; r[3] += g3; r[4] += g4;
        }
      }
    }
    // bottom terms: ubar' = -conj(PBhat) Bhat conj(That) Qbar Bhat conj(SBhat_{j+1})
    if (a.nb > 0) {
      const Q4<T> M = mul(mul(Bh, conj(Th)), mul(Qbar, Bh));
      Q4<T> S = bt.after;
      for (int jj = 0; jj < cb; jj++) {
        const int j = b0 + cb - 1 - jj;
        double g0 = 0, g1 = 0, g2 = 0, g3 = 0, g4 = 0;
        if (live && j < b1) {
          const RootRec64 rr = roots[a.nt + j];
          const Term<T> t = make_term(q, rr, T(rr.e));
          const Q4<T> right = S;
          S = mul(t.that, S);
          const Q4<T> P = (j > 0) ? mul(Bh, conj(S)) : Q4<T>{1, 0, 0, 0};
          const Q4<T> ub = sc(mul(mul(conj(P), M), conj(right)), T(-1));
          const Q4<T> bb = sc(pow_vjp(t, ub, T(rr.e)), Qmag);
          const T lv = t.rho > T(0) ? t.th / t.rho : T(0);
          const Q4<T> lg{t.lb, lv * t.b.x, lv * t.b.y, lv * t.b.z};
          const T ge = Qmag * dot(ub, mul(t.that, lg));
          qnew = Q4<T>{qnew.w + bb.w, qnew.x + bb.x, qnew.y + bb.y, qnew.z + bb.z};
          g0 = -(double)bb.w; g1 = -(double)bb.x; g2 = -(double)bb.y; g3 = -(double)bb.z; g4 = (double)ge;
          if ((jj & 7) == 7) S = unit(S);
        }
        g0 = across_groups<T>(g0);
        g1 = across_groups<T>(g1);
        g2 = across_groups<T>(g2);
        g3 = across_groups<T>(g3);
        g4 = across_groups<T>(g4);
        if (lane < kL && j < b1) {
          double* r = acc + wave * K + 5 * (a.nt + j);
          r[0] += g0; r[1] += g1; r[2] += g2; r[3] += g3; r[4] += g4;
        }
      }
    }
    qnew = group_sum_q(qnew);
    if (live) qbar = qnew;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < K; k += kAdjBlock) {
    double s = 0.0;
    for (int wv = 0; wv < kAdjBlock / 64; wv++) s += acc[wv * K + k];
    partials[(int64_t)blockIdx.x * K + k] = s;
  }
}

// sums the per-block partials and scatters: roots (n x 4) and powers (dJ/dp = s dJ/de)
__global__ void k_adjoint_finish(const double* __restrict__ partials, int64_t nb, int nt, int nr, double sc_top,
                                 double sc_bottom, double* __restrict__ grad_roots, double* __restrict__ grad_powers) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int K = 5 * nr;
  if (k >= K) return;
  double s = 0.0;
  for (int64_t b = 0; b < nb; b++) s += partials[b * K + k];
  const int i = k / 5, c = k - 5 * i;
  if (c < 4) grad_roots[4 * i + c] = s;
  else grad_powers[i] = s * (i < nt ? sc_top : sc_bottom);
}

qj_status points_adjoint(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec,
                         const double* d_pts, int64_t n, const double* d_dJdF, double* d_gr, double* d_gp,
                         hipStream_t s) {
  const int nr = a.nt + a.nb;
  const int K = 5 * nr;
  if (a.max_iter > kMaxAdjIter) {
    set_error("the adjoint supports max_iter <= %d (got %d)", kMaxAdjIter, a.max_iter);
    return QJ_ERR_CAPACITY;
  }
  for (const RootRec64& r : f->h64)
    if (!(r.e > 0.0)) {
      set_error("the adjoint needs positive exponents (power_scalar * power > 0)");
      return QJ_ERR_INVALID;
    }
  const size_t lds = sizeof(double) * (kAdjBlock / 64) * K;
  if (lds > 64 * 1024) {
    set_error("too many roots for the adjoint kernel (%d)", nr);
    return QJ_ERR_CAPACITY;
  }
  const int64_t per_block = kAdjBlock / (prec == QJ_FP64 ? adj_lanes<double>() : adj_lanes<float>());  // points
  const int64_t nblocks = n > 0 ? (n + per_block - 1) / per_block : 1;
  qj_status st = ensure_partials(ws, nblocks * K + K, s);
  if (st) return st;
  double* sums = ws->d_partials + nblocks * K;
  if (n > 0) {
    if (prec == QJ_FP64)
      hipLaunchKernelGGL(k_adjoint<double>, dim3((unsigned)nblocks), dim3(kAdjBlock), lds, s, a, f->d64, d_pts, n,
                         d_dJdF, ws->d_partials, K);
    else
      hipLaunchKernelGGL(k_adjoint<float>, dim3((unsigned)nblocks), dim3(kAdjBlock), lds, s, a, f->d64, d_pts, n,
                         d_dJdF, ws->d_partials, K);
    QJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_sum_partials, dim3(K), dim3(kSumThreads), 0, s, ws->d_partials, nblocks, K, sums);
    QJ_HIP(hipGetLastError());
  } else {
    QJ_HIP(hipMemsetAsync(sums, 0, sizeof(double) * K, s));
  }
  hipLaunchKernelGGL(k_adjoint_finish, dim3((K + 127) / 128), dim3(128), 0, s, sums, (int64_t)1, a.nt, nr,
                     a.sc_top, a.sc_bottom, d_gr, d_gp);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

}  // namespace qj
