// Julia field evaluation kernels for gfx950 (MI355X).
//
// Hot path: OPTIMIZE_3D::computeLogScaledPowerRationalMap (OPTIMIZE_3D.cpp:1272-1370)
// and its AP/cached variants. One voxel per lane; a 256-lane block covers a
// 16 x 4 x 4 (x,y,z) tile so a wave is a 16 x 4 x-major patch (neighbouring
// voxels escape together, stores coalesce in 128 B runs). Root records are
// read with uniform addresses inside the root loop, so they come through the
// scalar cache into SGPRs; no LDS is needed. There is no dense contraction
// here, so MFMA is not used: the work is VALU (FP64 FMA and the FP64 libm
// sequences, or FP32 FMA plus native transcendentals).
//
// Kernels:
//   k_grid64      FP64, reference expression order (the "plain FP64" method).
//   k_grid32      FP32 bulk with exact power-of-two range scaling; marks the
//                 band |F| < band, non-finite results and near-threshold
//                 escape / black-hole decisions into a compacted list
//                 (wave ballot + one atomic per block).
//   k_list64      FP64 re-evaluation of a compacted list (band or points).
//   k_list64x     FP64 with an explicit binary exponent (extended range) for
//                 cells that overflowed: the reference's long double re-pass.
#include <float.h>

#include "qj_geom.h"
#include "qj_internal.h"
#include "qj_quat.h"
#include "qj_eval.h"

namespace qj {

constexpr int kTileX = 16, kTileY = 4, kTileZ = 4, kBlock = 256;

// ---------------------------------------------------------------- geometry

struct GridSource {
  __device__ __forceinline__ static q64 start(const MapArgs& a, const double*, int64_t idx) {
    const int64_t sx = a.res[0], sxy = (int64_t)a.res[0] * a.res[1];
    const int z = (int)(idx / sxy);
    const int64_t r = idx - (int64_t)z * sxy;
    const int y = (int)(r / sx);
    const int x = (int)(r - (int64_t)y * sx);
    double c[3];
    cell_center(a, x, y, z + a.zoff, c);
    return {c[0], c[1], c[2], a.slice};
  }
};
struct PointSource {
  __device__ __forceinline__ static q64 start(const MapArgs& a, const double* pts, int64_t idx) {
    return {pts[3 * idx], pts[3 * idx + 1], pts[3 * idx + 2], a.slice};
  }
};

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum v over the block and add it to counter slot `slot` (spread copies).
__device__ __forceinline__ void block_count(unsigned long long v, unsigned long long* cnt, int slot) {
  __shared__ unsigned long long s_part[kBlock / 64];
  v = wave_sum_u64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_part[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += s_part[i];
    if (t) atomicAdd(&cnt[slot * kCounterSpread + (blockIdx.x % kCounterSpread)], t);
  }
  __syncthreads();
}

// Order-free compaction: flagged lanes append `value` to list; one atomic per block.
__device__ __forceinline__ void block_append(bool flag, uint32_t value, uint32_t* list, uint32_t* count) {
  __shared__ uint32_t s_cnt[kBlock / 64];
  __shared__ uint32_t s_base;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const unsigned long long m = __ballot(flag);
  const uint32_t rank = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_cnt[wave] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < kBlock / 64; i++) {
      const uint32_t c = s_cnt[i];
      s_cnt[i] = t;
      t += c;
    }
    s_base = t ? atomicAdd(count, t) : 0u;
  }
  __syncthreads();
  if (flag) list[s_base + s_cnt[wave] + rank] = value;
}

// ---------------------------------------------------------------- FP64 extended range

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
// QUATERNION::pow with exps = exp(e log|q|) kept as mantissa * 2^k.
__device__ __forceinline__ xq xpow(const q64& q, double e) {
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
  xq r{{em * c, scale2 * q.x, scale2 * q.y, scale2 * q.z}, k};
  xnorm(r);
  return r;
}
__device__ __forceinline__ xq xpoly(const RootRec64* __restrict__ r, int n, const q64& q) {
  xq res = xpow(qsub(q, q64{r[0].w, r[0].x, r[0].y, r[0].z}), r[0].e);
  for (int i = 1; i < n; i++) {
    const xq t = xpow(qsub(q, q64{r[i].w, r[i].x, r[i].y, r[i].z}), r[i].e);
    res.m = qmul(res.m, t.m);
    res.e += t.e;
    xnorm(res);
  }
  return res;
}
// computeRationalMapAP's long double re-pass (OPTIMIZE_3D.cpp:1496-1562) with
// FP64 mantissas and an integer binary exponent in place of x87 range.
__device__ __forceinline__ Res64 eval64x(const MapArgs& a, const RootRec64* __restrict__ roots, q64 q) {
  Res64 r{0.0, 0, true, false};
  double magnitude = qmag(q);
  if ((a.flags & QJ_MAP_SKIP_ORIGIN) && magnitude < 1e-8) {
    r.written = false;
    return r;
  }
  double logmag = log(magnitude);
  int it = 0;
  while (magnitude < a.escape && it < a.max_iter) {
    xq T = xpoly(roots, a.nt, q);
    xq v;
    if (a.nb > 0) {
      const xq B = xpoly(roots + a.nt, a.nb, q);
      v.m = qmul(T.m, qinv(B.m));
      v.e = T.e - B.e;
    } else {
      v = T;
    }
    v.m = qscale(v.m, a.exp_scaling);
    xnorm(v);
    const double mm = qmag(v.m);
    logmag = log(mm) + v.e * 0.6931471805599453;
    magnitude = ldexp(mm, v.e);
    q = q64{ldexp(v.m.w, v.e), ldexp(v.m.x, v.e), ldexp(v.m.y, v.e), ldexp(v.m.z, v.e)};
    it++;
    r.iters++;
    if ((a.flags & QJ_MAP_BLACK_HOLE) && magnitude < kBlackHole) it = a.max_iter;
  }
  r.F = isnan(logmag) ? 0.0 : logmag;  // NaN -> 0, OPTIMIZE_3D.cpp:1557-1561
  return r;
}

// ---------------------------------------------------------------- FP32 bulk


// Largest-component power-of-two renormalisation: exact, no rounding.
// Returns the applied exponent (0 for 0 / inf / nan).
__device__ __forceinline__ int renorm32(q32& t, int& E) {
  // max |component| in two instructions (fmaxf would canonicalise each input
  // first); a NaN component is dropped, and a NaN product is flagged by the
  // caller through its magnitude
  float mx;
  asm("v_max3_f32 %0, |%1|, |%2|, |%3|\n\tv_max_f32_e64 %0, %0, |%4|"
      : "=&v"(mx)
      : "v"(t.w), "v"(t.x), "v"(t.y), "v"(t.z));
  const int k = __builtin_amdgcn_frexp_expf(mx);
  t.w = __builtin_amdgcn_ldexpf(t.w, -k);
  t.x = __builtin_amdgcn_ldexpf(t.x, -k);
  t.y = __builtin_amdgcn_ldexpf(t.y, -k);
  t.z = __builtin_amdgcn_ldexpf(t.z, -k);
  E += k;
  return k;
}

// General-kind term (q - r)^e in FP32, following QUATERNION::pow's polar form
// with the magnitude |b|^e moved into the log2 accumulator L, so the
// quaternion part is unit length. The reference's guards carry over: v = 0
// gives a positive real result and a non-positive angle drops the imaginary
// part.
__device__ __forceinline__ q32 general32(float e, const q32& b, float& L) {
  const float partial = b.x * b.x + b.y * b.y + b.z * b.z;
  const float n2 = partial + b.w * b.w;
  const float vm = __builtin_sqrtf(partial);
  const float th = atan2f(vm, b.w);
  const float m = vm > 0.0f ? e * th : 0.0f;  // scale = e*acos(w/|q|)*0 when v = 0
  L += e * 0.5f * __log2f(n2);
  float s, c;
  __sincosf(m, &s, &c);
  const float f = (m > 0.0f && vm > 0.0f) ? s / vm : 0.0f;
  return q32{c, f * b.x, f * b.y, f * b.z};
}

// e = 1: pow(b, 1) is b itself except for v = 0, where the polar form gives
// +|w|. The FP32 path uses b and keeps the running minimum of |b.x| instead: a
// voxel that met b.x = 0 is sent to FP64, which applies the guard exactly. The
// minimum is one v_min_f32 with an |.| source modifier; fminf would add two
// canonicalising maxes per term. A NaN b.x leaves vx unchanged (the voxel is
// flagged through its magnitude).
__device__ __forceinline__ q32 kind1(const q32& b, float& vx) {
  asm("v_min_f32_e64 %0, %1, |%2|" : "=v"(vx) : "v"(vx), "v"(b.x));
  return b;
}

// e = 2: b * b (for v = 0 this is (w^2, 0, 0, 0), the polar form's value too).
__device__ __forceinline__ q32 kind2(const q32& b) {
  const float partial = b.x * b.x + b.y * b.y + b.z * b.z;
  const float t = 2.0f * b.w;
  return q32{b.w * b.w - partial, t * b.x, t * b.y, t * b.z};
}

// First term of a product; kb: the root's pattern bits (kPatSquare,
// kPatGeneral).
template <class EP>
__device__ __forceinline__ q32 term32(const F4& r, int kb, EP e, const q32& q, float& L,
                                      float& vx) {
  const q32 b = qsub(q, q32{r.w, r.x, r.y, r.z});
  if (kb & kPatGeneral) return general32(*e, b, L);
  if (kb & kPatSquare) return kind2(b);
  return kind1(b, vx);
}

// t <- t * (q - r)^e, with the product taken inside each kind's branch so the
// branches share no term registers (this file is compiled without
// SimplifyCFG's common-code sinking, which would merge the products behind
// register copies).
template <bool kGeneral, class EP>
__device__ __forceinline__ void apply32(q32& t, const F4& r, int kb, EP e, const q32& q,
                                        float& L, float& vx) {
  const q32 b = qsub(q, q32{r.w, r.x, r.y, r.z});
  if (kGeneral && (kb & kPatGeneral)) {
    t = qmul(t, general32(*e, b, L));
  } else if (kb & kPatSquare) {
    t = qmul(t, kind2(b));
  } else {
    t = qmul(t, kind1(b, vx));
  }
}

// Ordered product over n roots with exponent tracking, renormalised every kRN
// terms. hi/lo receive the largest and smallest log2 magnitude of the running
// product at the renormalisation points (E, plus L when general kinds exist);
// `bad` is set when a group of kRN terms fell below 2^-60 (kRN = 4 only: a
// group of four can leave FP32's normal range, two cannot). vx: minimum |x| over
// the e = 1 terms (see kind1). The roots of a group are read before its first
// term: one scalar load of kRN * 16 bytes and one pattern word.
template <class QP>
__device__ __forceinline__ F4 load_root(QP p) {
  return F4{p->w, p->x, p->y, p->z};
}

template <int kRN, bool kGeneral, class QP, class PP, class EP>
__device__ __forceinline__ q32 poly32(QP rq, PP rp, EP re, int n, const q32& q, int& E, float& L, float& hi,
                                      float& lo, bool& bad, float& vx) {
  q32 t = term32(load_root(rq), rp[0], re, q, L, vx);
  int i = 1;
  int mn = 0;
  int ehi = -(1 << 30), elo = 1 << 30;
  for (; i + kRN - 1 < n; i += kRN) {
    F4 g[kRN];
#pragma unroll
    for (int j = 0; j < kRN; j++) g[j] = load_root(rq + i + j);
    const int pat = rp[i];
#pragma unroll
    for (int j = 0; j < kRN; j++) apply32<kGeneral>(t, g[j], pat >> (2 * j), re + i + j, q, L, vx);
    const int k = renorm32(t, E);
    if (kRN > 2) mn = min(mn, k);
    if (kGeneral) {
      const float l = (float)E + L;
      hi = fmaxf(hi, l);
      lo = fminf(lo, l);
    } else {
      ehi = max(ehi, E);
      elo = min(elo, E);
    }
  }
  for (; i < n; i++) apply32<kGeneral>(t, load_root(rq + i), rp[i], re + i, q, L, vx);
  const int k = renorm32(t, E);
  if (kRN > 2) bad |= min(mn, k) < -60;
  if (kGeneral) {
    const float l = (float)E + L;
    hi = fmaxf(hi, l);
    lo = fminf(lo, l);
  } else {
    ehi = max(ehi, E);
    elo = min(elo, E);
    hi = fmaxf(hi, (float)ehi);
    lo = fminf(lo, (float)elo);
  }
  return t;
}

// Renormalising every 4 terms needs 4 terms of up to 2^term_bits to stay in
// FP32 range: 4 * 30 + 1 < 127.
constexpr double kRN4MaxTermBits = 30.0;

struct Res32 {
  float F;
  int iters;
  bool written;
  bool need64;
};

// One voxel in FP32 with the FP64 method's range behaviour reproduced where it
// is certain: every running product of T and B must stay well inside FP64's
// range (term_bits covers the product between two renormalisations); then
// |B| > 2^512.5 means |B|^2 overflows in FP64, B^-1 = conj(B) * 0 and q = 0
// exactly (F = -inf under the black-hole stop); anything closer to a range
// edge, or outside it, goes to FP64.
template <int kRN, bool kGeneral, class RootsT>
__device__ __forceinline__ Res32 eval32(const MapArgs& a, const RootsT& roots, const Consts32& k,
                                        const q64& q0) {
  Res32 res{0.0f, 0, true, false};
  // the first escape / origin test in FP64, exactly as the reference
  const double mag0 = qmag(q0);
  if ((a.flags & QJ_MAP_SKIP_ORIGIN) && mag0 < 1e-8) {
    res.written = false;
    return res;
  }
  if (!(mag0 < a.escape) || a.max_iter <= 0) {
    res.F = (float)log(mag0);
    res.need64 = true;  // trivial case: let FP64 write it exactly
    return res;
  }
  const bool black_hole = (a.flags & QJ_MAP_BLACK_HOLE) != 0;
  q32 q{(float)q0.w, (float)q0.x, (float)q0.y, (float)q0.z};
  float log2mag = 0.0f;
  float vx = INFINITY;
  bool amb = false, bad = false, zero = false;
  int it = 0;
  // FP64 running products between two checkpoints grow by at most (kRN - 1) terms
  const float lim = 1020.0f - (float)(kRN - 1) * k.term_bits;
  while (true) {
    int ET = 0, EB = 0;
    float L = 0.0f, hiT = -1e30f, loT = 1e30f;
    const q32 T = poly32<kRN, kGeneral>(roots.q, roots.pat, roots.e, a.nt, q, ET, L, hiT, loT, bad, vx);
    bad |= !(hiT < lim && loT > -lim);
    float l2 = k.log2c;
    if (a.nb > 0) {
      float LB = 0.0f, hiB = -1e30f, loB = 1e30f;
      const q32 B = poly32<kRN, kGeneral>(roots.q + a.nt, roots.pat + a.nt, roots.e + a.nt, a.nb, q, EB, LB, hiB,
                                          loB, bad, vx);
      bad |= !(hiB < lim && loB > -lim);
      const float nb2 = B.w * B.w + B.x * B.x + B.y * B.y + B.z * B.z;
      const float lB = 0.5f * __log2f(nb2) + (float)EB + LB;
      const q32 D = qmul(T, q32{B.w, -B.x, -B.y, -B.z});  // T conj(B) = T B^-1 |B|^2
      const float nd2 = D.w * D.w + D.x * D.x + D.y * D.y + D.z * D.z;
      l2 += 0.5f * __log2f(nd2) - __log2f(nb2) + (float)(ET - EB) + (L - LB);
      bad |= !(nb2 > 0.0f) || !(nd2 > 0.0f);
      if (lB > 512.5f) {
        zero = true;  // FP64: |B|^2 = inf, 1/inf = 0, q = T * 0 = 0
      } else if (lB > 511.5f || lB < -510.0f) {
        bad = true;   // within FP32 error of the FP64 overflow edge of |B|^2 or 1/|B|^2
      } else if (!bad) {
        const float sc = exp2f(k.log2c + (float)(ET - EB) + (L - LB)) / nb2;
        q = qscale(D, sc);
      }
    } else {
      const float nt2 = T.w * T.w + T.x * T.x + T.y * T.y + T.z * T.z;
      l2 += 0.5f * __log2f(nt2) + (float)ET + L;
      bad |= !(nt2 > 0.0f);
      if (!bad) q = qscale(T, exp2f(k.log2c + (float)ET + L));
    }
    it++;
    res.iters++;
    if (zero) {
      // q = 0: magnitude 0 < 10 FLT_MIN stops the loop with F = log(0)
      bad |= !black_hole;
      log2mag = -INFINITY;
      break;
    }
    log2mag = l2;
    bad |= !isfinite(l2) || l2 > 511.0f || l2 < -530.0f;  // |q|^2 over/underflow in FP64
    amb |= fabsf(l2 - k.log2esc) < k.margin;
    const bool bh = black_hole && l2 < k.log2bh;
    if (black_hole) amb |= fabsf(l2 - k.log2bh) < k.margin;
    if (bad || l2 >= k.log2esc || bh || it >= a.max_iter) break;
  }
  res.F = zero ? -INFINITY : log2mag * 0.69314718f;
  res.need64 = bad || amb || !(vx > 0.0f) || !(fabsf(res.F) >= k.band);
  return res;
}

// ---------------------------------------------------------------- kernels

__device__ __forceinline__ bool tile_voxel(const MapArgs& a, int& x, int& y, int& z) {
  const int ntx = (a.res[0] + kTileX - 1) / kTileX, nty = (a.res[1] + kTileY - 1) / kTileY;
  const int b = blockIdx.x;
  const int tx = b % ntx, ty = (b / ntx) % nty, tz = b / (ntx * nty);
  const int t = threadIdx.x;
  x = tx * kTileX + (t & (kTileX - 1));
  y = ty * kTileY + ((t >> 4) & (kTileY - 1));
  z = tz * kTileZ + (t >> 6);
  return x < a.res[0] && y < a.res[1] && z < a.res[2];
}

__global__ __launch_bounds__(kBlock) void k_grid64(MapArgs a, const RootRec64* __restrict__ roots,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   uint32_t* __restrict__ alist, uint32_t* __restrict__ acount) {
  int x, y, z;
  const bool valid = tile_voxel(a, x, y, z);
  int iters = 0;
  bool ovf = false, nonfinite = false;
  const int64_t idx = x + (int64_t)y * a.res[0] + (int64_t)z * a.res[0] * a.res[1];
  if (valid) {
    double c[3];
    cell_center(a, x, y, z + a.zoff, c);
    const Res64 r = eval64(a, roots, q64{c[0], c[1], c[2], a.slice});
    if (r.written) {
      out[idx] = r.F;
      nonfinite = !isfinite(r.F);
    }
    iters = r.iters;
    ovf = r.overflow;
  }
  block_count(valid ? 1 : 0, cnt, CNT_VOXELS);
  block_count(iters, cnt, CNT_ITERS);
  block_count(nonfinite ? 1 : 0, cnt, CNT_NONFINITE);
  if (a.flags & QJ_MAP_ADAPTIVE) {
    block_append(ovf, (uint32_t)idx, alist, acount);
    block_count(ovf ? 1 : 0, cnt, CNT_ADAPTIVE);
  }
}

template <int kRN, bool kGeneral>
__global__ __launch_bounds__(kBlock) void k_grid32(MapArgs a, Consts32 k, const F4* __restrict__ roots,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount) {
  const Roots32 R = roots32_view(roots, a.nt + a.nb);
  int x, y, z;
  const bool valid = tile_voxel(a, x, y, z);
  int iters = 0;
  bool need = false;
  const int64_t idx = x + (int64_t)y * a.res[0] + (int64_t)z * a.res[0] * a.res[1];
  if (valid) {
    double c[3];
    cell_center(a, x, y, z + a.zoff, c);
    const Res32 r = eval32<kRN, kGeneral>(a, R, k, q64{c[0], c[1], c[2], a.slice});
    if (r.written) {
      if (!r.need64) out[idx] = (double)r.F;
      need = r.need64;
    }
    iters = r.iters;
  }
  block_count(valid ? 1 : 0, cnt, CNT_VOXELS);
  block_count(iters, cnt, CNT_ITERS);
  block_append(need, (uint32_t)idx, blist, bcount);
}

template <class Src, int kRN, bool kGeneral>
__global__ __launch_bounds__(kBlock) void k_points32(MapArgs a, Consts32 k, const F4* __restrict__ roots,
                                                     const double* __restrict__ pts, int64_t n,
                                                     double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                     uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount) {
  const Roots32 R = roots32_view(roots, a.nt + a.nb);
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  int iters = 0;
  bool need = false;
  if (valid) {
    const Res32 r = eval32<kRN, kGeneral>(a, R, k, Src::start(a, pts, i));
    if (r.written) {
      if (!r.need64) out[i] = (double)r.F;
      need = r.need64;
    } else {
      out[i] = __builtin_nan("");
    }
    iters = r.iters;
  }
  block_count(valid ? 1 : 0, cnt, CNT_VOXELS);
  block_count(iters, cnt, CNT_ITERS);
  block_append(need, (uint32_t)i, blist, bcount);
}

// FP64 over a compacted list (count read on the device), or over all n
// indices when list == nullptr. kPow: the PowMode (kExactInt / kExactOnly for
// the FP32 path's band and range re-evaluation).
template <class Src, bool kExtended, int kPow = kPolar>
__global__ __launch_bounds__(kBlock) void k_list64(MapArgs a, const RootRec64* __restrict__ roots,
                                                   const double* __restrict__ pts,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count, int64_t n_all,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   int slot_vox, int slot_iter, uint32_t* __restrict__ alist,
                                                   uint32_t* __restrict__ acount) {
  const int64_t n = list ? (int64_t)*count : n_all;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  unsigned long long vox = 0, iters = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    bool ovf = false;
    uint32_t idx = 0;
    if (i < n) {
      idx = list ? list[i] : (uint32_t)i;
      const q64 q0 = Src::start(a, pts, idx);
      const Res64 r = kExtended ? eval64x(a, roots, q0) : eval64<kPow>(a, roots, q0);
      if (r.written) out[idx] = r.F;
      else if (pts) out[idx] = __builtin_nan("");
      vox++;
      iters += r.iters;
      ovf = r.overflow;
    }
    if (!kExtended && (a.flags & QJ_MAP_ADAPTIVE)) block_append(ovf, idx, alist, acount);
  }
  block_count(vox, cnt, slot_vox);
  block_count(iters, cnt, slot_iter);
}

// ---------------------------------------------------------------- batched chunks (request loop)

__global__ void k_zero_u32(uint32_t* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0u;
}

__device__ __forceinline__ q64 chunk_start(const MapArgs& a, int64_t idx) { return GridSource::start(a, nullptr, idx); }

// Work queue over (chunk, block-of-chunk) tiles: a resident grid takes tiles
// with one atomic per tile until every tile of its chunk range is taken, so a
// launch costs what its used chunks cost (no empty blocks) and blocks that
// draw cheap tiles take more of them. Every block leaves when the queue passes
// the tile count.
__device__ __forceinline__ int next_tile(uint32_t* queue) {
  __shared__ int s_tile;
  __syncthreads();  // the previous tile's readers are done with s_tile
  if (threadIdx.x == 0) s_tile = (int)atomicAdd(queue, 1u);
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(s_tile);  // uniform: descriptor reads become scalar loads
}

// A root block read through the constant address space: its loads are known
// not to be clobbered by the kernel's stores, so a block-uniform address gives
// scalar loads even when the pointer itself was read from memory.
struct Roots32K {
  const __attribute__((address_space(4))) F4* q;
  const __attribute__((address_space(4))) int32_t* pat;
  const __attribute__((address_space(4))) float* e;
};
__device__ __forceinline__ Roots32K roots32_const(const F4* blk, int n) {
  const Roots32 r = roots32_view(blk, n);
  return Roots32K{(const __attribute__((address_space(4))) F4*)r.q,
                  (const __attribute__((address_space(4))) int32_t*)r.pat,
                  (const __attribute__((address_space(4))) float*)r.e};
}

// Batch counters (uint32): [0] band count of integer-exponent chunks (list from
// the front), [1] band count of general chunks (list from the back), [2..4]
// FP32 queues per variant, [5] FP64 queue.
enum { BC_BAND_INT = 0, BC_BAND_GEN = 1, BC_Q32 = 2, BC_Q64 = 5, BC_N = 8 };

// Descriptors are grouped by FP32 variant: chunks [0, g1) have general
// exponents, [g1, g2) integer exponents renormalised every 4 terms, [g2,
// total) integer exponents renormalised every 2.
template <int kVar>
__global__ __launch_bounds__(kBlock) void k_batch32(const ChunkDesc* __restrict__ descs, int blocks_per_chunk,
                                                    int chunk_voxels, uint32_t* __restrict__ blist, int64_t list_cap,
                                                    uint32_t* __restrict__ counters) {
  constexpr int kRN = kVar == 1 ? 4 : 2;
  constexpr bool kGen = kVar == 0;
  const ChunkDesc& d0 = descs[0];
  const int c0 = kVar == 0 ? 0 : (kVar == 1 ? d0.g1 : d0.g2);
  const int c1 = kVar == 0 ? d0.g1 : (kVar == 1 ? d0.g2 : d0.total);
  const int tiles = (c1 - c0) * blocks_per_chunk;
  for (int tile = next_tile(counters + BC_Q32 + kVar); tile < tiles; tile = next_tile(counters + BC_Q32 + kVar)) {
    const int c = c0 + tile / blocks_per_chunk;
    const int local = (tile % blocks_per_chunk) * kBlock + threadIdx.x;
    const ChunkDesc& d = descs[c];
    bool need = false;
    if (local < d.count) {
      const MapArgs& a = d.a;
      const Consts32 k = d.k;
      const int64_t idx = d.start + local;
      const Res32 r = eval32<kRN, kGen>(a, roots32_const(d.r32, a.nt + a.nb), k, chunk_start(a, idx));
      if (r.written) {
        if (!r.need64) d.out[idx] = (double)r.F;
        need = r.need64;
      }
    }
    const uint32_t e = (uint32_t)(c * chunk_voxels + local);
    if (kGen) {
      // general chunks fill the list from the back: entry k at list_cap - 1 - k
      __shared__ uint32_t s_gbase;
      const unsigned long long m = __ballot(need);
      __shared__ uint32_t s_gcnt[kBlock / 64];
      const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
      if (lane == 0) s_gcnt[wave] = (uint32_t)__popcll(m);
      __syncthreads();
      if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kBlock / 64; w++) {
          const uint32_t cw = s_gcnt[w];
          s_gcnt[w] = t;
          t += cw;
        }
        s_gbase = t ? atomicAdd(counters + BC_BAND_GEN, t) : 0u;
      }
      __syncthreads();
      if (need) {
        const uint32_t k = s_gbase + s_gcnt[wave] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        blist[list_cap - 1 - (int64_t)k] = e;
      }
    } else {
      block_append(need, e, blist, counters + BC_BAND_INT);
    }
  }
}

// FP64 for the band list of one exponent class (integer exponents as exact
// products); kFromEnd reads the general chunks' list from the back.
template <int kPow, bool kFromEnd>
__global__ __launch_bounds__(kBlock) void k_batch64_list(const ChunkDesc* __restrict__ descs, int chunk_voxels,
                                                         const uint32_t* __restrict__ list, int64_t list_cap,
                                                         const uint32_t* __restrict__ count) {
  const int64_t n = *count;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    const uint32_t e = kFromEnd ? list[list_cap - 1 - i] : list[i];
    const int c = (int)(e / (uint32_t)chunk_voxels);
    const int local = (int)(e - (uint32_t)c * (uint32_t)chunk_voxels);
    // consecutive entries nearly always come from one chunk: then the chunk
    // index is wave-uniform and its root records come in scalar loads
    const int cw = __builtin_amdgcn_readfirstlane(c);
    if (__ballot(c != cw) == 0ull) {
      const ChunkDesc& d = descs[cw];
      const int64_t idx = d.start + local;
      typedef const __attribute__((address_space(4))) RootRec64 cR64;
      const Res64 r = eval64<kPow>(d.a, (cR64*)d.r64, chunk_start(d.a, idx));
      if (r.written) d.out[idx] = r.F;
    } else {
      const ChunkDesc& d = descs[c];
      const int64_t idx = d.start + local;
      const Res64 r = eval64<kPow>(d.a, d.r64, chunk_start(d.a, idx));
      if (r.written) d.out[idx] = r.F;
    }
  }
}

// FP64 (the reference method) for every voxel of every chunk, work-queue tiles
__global__ __launch_bounds__(kBlock) void k_batch64(const ChunkDesc* __restrict__ descs, int blocks_per_chunk,
                                                    uint32_t* __restrict__ counters) {
  const int tiles = descs[0].total * blocks_per_chunk;
  for (int tile = next_tile(counters + BC_Q64); tile < tiles; tile = next_tile(counters + BC_Q64)) {
    const int c = tile / blocks_per_chunk;
    const int local = (tile - c * blocks_per_chunk) * kBlock + threadIdx.x;
    const ChunkDesc& d = descs[c];
    if (local < d.count) {
      const int64_t idx = d.start + local;
      const Res64 r = eval64(d.a, d.r64, chunk_start(d.a, idx));
      if (r.written) d.out[idx] = r.F;
    }
  }
}

static int resident_blocks(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2048;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
    return 2 * cus;
  return per_cu * cus;
}

int batch_variant(const MapArgs& a) { return a.gen ? 0 : (a.term_bits <= kRN4MaxTermBits ? 1 : 2); }

qj_status launch_batch(const ChunkDesc* d_desc, int nchunks, int chunk_voxels, qj_precision prec, uint32_t* d_band,
                       uint32_t* d_counters, unsigned variants, hipStream_t s) {
  const int bpc = (chunk_voxels + kBlock - 1) / kBlock;
  const int64_t cap = (int64_t)nchunks * chunk_voxels;
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, d_counters, (int)BC_N);
  QJ_HIP(hipGetLastError());
  if (prec == QJ_FP64) {
    static const int g64 = resident_blocks((const void*)k_batch64);
    hipLaunchKernelGGL(k_batch64, dim3(g64), dim3(kBlock), 0, s, d_desc, bpc, d_counters);
    QJ_HIP(hipGetLastError());
    return QJ_OK;
  }
  static const int g0 = resident_blocks((const void*)k_batch32<0>);
  static const int g1 = resident_blocks((const void*)k_batch32<1>);
  static const int g2 = resident_blocks((const void*)k_batch32<2>);
  // a variant with no chunks is not launched: an empty persistent grid still waits for free
  // compute units behind the other in-flight launch
  if (variants & 2u) {
    hipLaunchKernelGGL(k_batch32<1>, dim3(g1), dim3(kBlock), 0, s, d_desc, bpc, chunk_voxels, d_band, cap, d_counters);
    QJ_HIP(hipGetLastError());
  }
  if (variants & 4u) {
    hipLaunchKernelGGL(k_batch32<2>, dim3(g2), dim3(kBlock), 0, s, d_desc, bpc, chunk_voxels, d_band, cap, d_counters);
    QJ_HIP(hipGetLastError());
  }
  if (variants & 1u) {
    hipLaunchKernelGGL(k_batch32<0>, dim3(g0), dim3(kBlock), 0, s, d_desc, bpc, chunk_voxels, d_band, cap, d_counters);
    QJ_HIP(hipGetLastError());
  }
  if (variants & 6u) {  // band voxels of the integer-exponent variants
    hipLaunchKernelGGL((k_batch64_list<kExactOnly, false>), dim3(2048), dim3(kBlock), 0, s, d_desc, chunk_voxels,
                       (const uint32_t*)d_band, cap, (const uint32_t*)(d_counters + BC_BAND_INT));
    QJ_HIP(hipGetLastError());
  }
  if (variants & 1u) {  // band voxels of the general variant
    hipLaunchKernelGGL((k_batch64_list<kExactInt, true>), dim3(1024), dim3(kBlock), 0, s, d_desc, chunk_voxels,
                       (const uint32_t*)d_band, cap, (const uint32_t*)(d_counters + BC_BAND_GEN));
    QJ_HIP(hipGetLastError());
  }
  return QJ_OK;
}

// ---------------------------------------------------------------- launchers

Consts32 consts32(const MapArgs& a, double band) {
  Consts32 k;
  k.log2c = (float)log2(a.exp_scaling);
  k.log2esc = (float)log2(a.escape);
  k.log2bh = (float)log2(kBlackHole);
  k.margin = 0.02f;
  k.band = (float)band;
  k.term_bits = (float)a.term_bits;
  return k;
}

static int list_grid(int64_t n_hint) {
  const int64_t b = (n_hint + kBlock - 1) / kBlock;
  return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

static qj_status run_adaptive(qj_workspace* ws, const MapArgs& a, const qj_rational* f, const double* pts,
                              int64_t n, double* d_out, hipStream_t s) {
  if (!(a.flags & QJ_MAP_ADAPTIVE)) return QJ_OK;
  uint32_t* alist = ws->d_list + ws->list_cap;
  if (pts)
    hipLaunchKernelGGL((k_list64<PointSource, true>), dim3(list_grid(n / 64 + 1)), dim3(kBlock), 0, s, a, f->d64,
                       pts, alist, ws->d_list_count + 1, (int64_t)0, d_out, ws->d_counters, (int)CNT_BAND_VOXELS,
                       (int)CNT_BAND_ITERS, (uint32_t*)nullptr, (uint32_t*)nullptr);
  else
    hipLaunchKernelGGL((k_list64<GridSource, true>), dim3(list_grid(n / 64 + 1)), dim3(kBlock), 0, s, a, f->d64,
                       pts, alist, ws->d_list_count + 1, (int64_t)0, d_out, ws->d_counters, (int)CNT_BAND_VOXELS,
                       (int)CNT_BAND_ITERS, (uint32_t*)nullptr, (uint32_t*)nullptr);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

qj_status launch_field(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec, double band,
                       double* d_out, hipStream_t s) {
  const int64_t nvox = (int64_t)a.res[0] * a.res[1] * a.res[2];
  if (nvox <= 0) return QJ_OK;
  if (nvox > (int64_t)UINT32_MAX) {
    set_error("grid of %lld cells exceeds 2^32; evaluate it in tiles", (long long)nvox);
    return QJ_ERR_CAPACITY;
  }
  qj_status st = ensure_list(ws, nvox, s);
  if (st) return st;
  // list counters cleared by a kernel, not a memset node, so the call can be captured in a graph
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, ws->d_list_count, 2);
  QJ_HIP(hipGetLastError());
  const int64_t ntx = (a.res[0] + kTileX - 1) / kTileX, nty = (a.res[1] + kTileY - 1) / kTileY,
                ntz = (a.res[2] + kTileZ - 1) / kTileZ;
  const int64_t nblocks = ntx * nty * ntz;
  if (nblocks > INT32_MAX) return QJ_ERR_CAPACITY;
  uint32_t* alist = ws->d_list + ws->list_cap;
  timing_mark(ws, s);
  if (prec == QJ_FP64) {
    hipLaunchKernelGGL(k_grid64, dim3((unsigned)nblocks), dim3(kBlock), 0, s, a, f->d64, d_out, ws->d_counters,
                       alist, ws->d_list_count + 1);
    QJ_HIP(hipGetLastError());
    timing_mark(ws, s);
    timing_mark(ws, s);
  } else {
    const Consts32 k = consts32(a, band);
    const dim3 g((unsigned)nblocks), b(kBlock);
    if (a.gen)
      hipLaunchKernelGGL((k_grid32<2, true>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, ws->d_counters, ws->d_list,
                         ws->d_list_count);
    else if (a.term_bits <= kRN4MaxTermBits)
      hipLaunchKernelGGL((k_grid32<4, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, ws->d_counters, ws->d_list,
                         ws->d_list_count);
    else
      hipLaunchKernelGGL((k_grid32<2, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, ws->d_counters, ws->d_list,
                         ws->d_list_count);
    QJ_HIP(hipGetLastError());
    timing_mark(ws, s);
    timing_mark(ws, s);
    if (a.gen)
      hipLaunchKernelGGL((k_list64<GridSource, false, kExactInt>), dim3(list_grid(nvox / 16)), dim3(kBlock), 0, s, a,
                         f->d64, (const double*)nullptr, ws->d_list, ws->d_list_count, (int64_t)0, d_out,
                         ws->d_counters, (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, alist, ws->d_list_count + 1);
    else
      hipLaunchKernelGGL((k_list64<GridSource, false, kExactOnly>), dim3(list_grid(nvox / 16)), dim3(kBlock), 0, s, a,
                         f->d64, (const double*)nullptr, ws->d_list, ws->d_list_count, (int64_t)0, d_out,
                         ws->d_counters, (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, alist, ws->d_list_count + 1);
    QJ_HIP(hipGetLastError());
  }
  timing_mark(ws, s);
  return run_adaptive(ws, a, f, nullptr, nvox, d_out, s);
}

qj_status launch_points(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec, double band,
                        const double* d_pts, int64_t n, double* d_out, hipStream_t s) {
  if (n <= 0) return QJ_OK;
  if (n > (int64_t)UINT32_MAX) return QJ_ERR_CAPACITY;
  qj_status st = ensure_list(ws, n, s);
  if (st) return st;
  // list counters cleared by a kernel, not a memset node, so the call can be captured in a graph
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, ws->d_list_count, 2);
  QJ_HIP(hipGetLastError());
  uint32_t* alist = ws->d_list + ws->list_cap;
  if (prec == QJ_FP64) {
    hipLaunchKernelGGL((k_list64<PointSource, false>), dim3(list_grid(n)), dim3(kBlock), 0, s, a, f->d64, d_pts,
                       (const uint32_t*)nullptr, (const uint32_t*)nullptr, n, d_out, ws->d_counters,
                       (int)CNT_VOXELS, (int)CNT_ITERS, alist, ws->d_list_count + 1);
    QJ_HIP(hipGetLastError());
  } else {
    const int64_t nb = (n + kBlock - 1) / kBlock;
    const Consts32 k = consts32(a, band);
    const dim3 g((unsigned)nb), b(kBlock);
    if (a.gen)
      hipLaunchKernelGGL((k_points32<PointSource, 2, true>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         ws->d_counters, ws->d_list, ws->d_list_count);
    else if (a.term_bits <= kRN4MaxTermBits)
      hipLaunchKernelGGL((k_points32<PointSource, 4, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         ws->d_counters, ws->d_list, ws->d_list_count);
    else
      hipLaunchKernelGGL((k_points32<PointSource, 2, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         ws->d_counters, ws->d_list, ws->d_list_count);
    QJ_HIP(hipGetLastError());
    if (a.gen)
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactInt>), dim3(list_grid(n / 16)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, ws->d_list, ws->d_list_count, (int64_t)0, d_out, ws->d_counters,
                         (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, alist, ws->d_list_count + 1);
    else
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactOnly>), dim3(list_grid(n / 16)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, ws->d_list, ws->d_list_count, (int64_t)0, d_out, ws->d_counters,
                         (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, alist, ws->d_list_count + 1);
    QJ_HIP(hipGetLastError());
  }
  return run_adaptive(ws, a, f, d_pts, n, d_out, s);
}

}  // namespace qj
