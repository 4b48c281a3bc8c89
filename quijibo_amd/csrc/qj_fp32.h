// FP32 bulk evaluation of the map (the device code of k_grid32, k_points32
// and k_batch32). Hot path: OPTIMIZE_3D::computeLogScaledPowerRationalMap
// (OPTIMIZE_3D.cpp:1272-1370) in FP32 with exact power-of-two range scaling.
#pragma once
#include "qj_geom.h"
#include "qj_quat.h"
#include "qj_types.h"

namespace qj {

constexpr int kTileX = 16, kTileY = 4, kTileZ = 4, kBlock = 256;

// ---------------------------------------------------------------- wave helpers

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum v over the block and add it to counter slot `slot` (kSpread copies per
// slot, one chosen by the block index, to spread the atomics).
template <int kSpread = kCounterSpread>
__device__ __forceinline__ void block_count(unsigned long long v, unsigned long long* cnt, int slot) {
  __shared__ unsigned long long s_part[kBlock / 64];
  v = wave_sum_u64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_part[wave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int i = 0; i < kBlock / 64; i++) t += s_part[i];
    if (t) atomicAdd(&cnt[slot * kSpread + (blockIdx.x % kSpread)], t);
  }
  __syncthreads();
}

// Order-free compaction: flagged lanes append `value` to list; one atomic per
// block. Entries at `cap` and beyond are not written (the count still counts
// them); returns whether this lane's entry was written.
__device__ __forceinline__ bool block_append(bool flag, uint32_t value, uint32_t* list, uint32_t* count,
                                             uint32_t cap) {
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
  const uint32_t slot = s_base + s_cnt[wave] + rank;
  const bool stored = flag && slot < cap;
  if (stored) list[slot] = value;
  return stored;
}

// What the FP32 pass writes into a cell that needs FP64 but found the band
// list full (a signalling-NaN pattern no arithmetic produces); the band sweep
// evaluates those cells.
constexpr unsigned long long kBandMark = 0x7ff40000b0a4d5eeull;

// ---------------------------------------------------------------- FP32 bulk


// The quaternion product and q - r with every rounding fixed by the source:
// each product is an explicit fused multiply-add and contraction is off for
// the rest. The compiler's own contraction choices depend on how many uses a
// product has, so the same source could round differently in the grid, point
// and batch kernels, or after an unrelated change, where the roots make common
// subexpressions visible (a root with a zero component gives q - r = q).
__device__ __forceinline__ q32 qmul32(const q32& l, const q32& r) {
#pragma clang fp contract(off)
  q32 f;
  f.x = __builtin_fmaf(l.w, r.x, __builtin_fmaf(r.w, l.x, __builtin_fmaf(-l.z, r.y, l.y * r.z)));
  f.y = __builtin_fmaf(l.w, r.y, __builtin_fmaf(r.w, l.y, __builtin_fmaf(-l.x, r.z, l.z * r.x)));
  f.z = __builtin_fmaf(l.w, r.z, __builtin_fmaf(r.w, l.z, __builtin_fmaf(-l.y, r.x, l.x * r.y)));
  f.w = __builtin_fmaf(-l.z, r.z, __builtin_fmaf(-r.y, l.y, __builtin_fmaf(-l.x, r.x, l.w * r.w)));
  return f;
}
__device__ __forceinline__ q32 qsub32(const q32& a, const F4& r) {
#pragma clang fp contract(off)
  return q32{a.w - r.w, a.x - r.x, a.y - r.y, a.z - r.z};
}

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
// The angle of w + i y for y >= 0, in [0, pi], with finite inputs (|q| is
// below the escape radius). Octant reduction to t = min(|w|, y) / max(|w|, y)
// in [0, 1], then atan(t) = t P(t^2), P of degree 8 fitted to atan(t)/t by
// least squares (relative weight) at 4000 Chebyshev nodes of t^2 in [0, 1]:
// at most 1.7 ulp from the correctly rounded atan over [0, 1] in FP32 Horner
// form, against about 40 instructions of the library atan2f's special cases.
// y = w = 0 gives NaN; the caller does not use the angle when y = 0. The
// octant test is one compare shared by the two selects and the final fold
// (no inline asm here: y comes straight from a transcendental instruction,
// and only the compiler's own instructions get the wait such a read needs).
__device__ __forceinline__ float angle_upper(float y, float w) {
  const float aw = fabsf(w);
  const bool steep = y > aw;
  const float mx = steep ? y : aw, mn = steep ? aw : y;
  const float t = mn * __builtin_amdgcn_rcpf(mx);
  const float s = t * t;
  float p = 0.0027817077934741974f;
  p = __builtin_fmaf(p, s, -0.01579383574426174f);
  p = __builtin_fmaf(p, s, 0.042241036891937256f);
  p = __builtin_fmaf(p, s, -0.07465779036283493f);
  p = __builtin_fmaf(p, s, 0.10622653365135193f);
  p = __builtin_fmaf(p, s, -0.14198920130729675f);
  p = __builtin_fmaf(p, s, 0.1999201774597168f);
  p = __builtin_fmaf(p, s, -0.3333304524421692f);
  p = __builtin_fmaf(p, s, 1.0f);
  float r = t * p;
  r = steep ? 1.5707963267948966f - r : r;
  return w < 0.0f ? 3.141592653589793f - r : r;
}

__device__ __forceinline__ q32 general32(float e, const q32& b, float& L) {
#pragma clang fp contract(off)
  const float partial = __builtin_fmaf(b.z, b.z, __builtin_fmaf(b.y, b.y, b.x * b.x));
  const float n2 = __builtin_fmaf(b.w, b.w, partial);
  const float vm = __builtin_amdgcn_sqrtf(partial);  // within 1 ulp; the band repeats close cells in FP64
  const float th = angle_upper(vm, b.w);
  const float m = vm > 0.0f ? e * th : 0.0f;  // scale = e*acos(w/|q|)*0 when v = 0
  L = __builtin_fmaf(e * 0.5f, __log2f(n2), L);
  float s, c;
  __sincosf(m, &s, &c);
  const float f = (m > 0.0f && vm > 0.0f) ? s * __builtin_amdgcn_rcpf(vm) : 0.0f;
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
#pragma clang fp contract(off)
  const float partial = __builtin_fmaf(b.z, b.z, __builtin_fmaf(b.y, b.y, b.x * b.x));
  const float t = 2.0f * b.w;
  return q32{__builtin_fmaf(b.w, b.w, -partial), t * b.x, t * b.y, t * b.z};
}

// First term of a product; kb: the root's pattern bits (kPatSquare,
// kPatGeneral); r.ex(i) is read only for a general exponent.
template <bool kGeneral, class Acc>
__device__ __forceinline__ q32 term32(const F4& rt, int kb, const Acc& r, int i, const q32& q, float& L, float& vx) {
  const q32 b = qsub32(q, rt);
  if (kGeneral && (kb & kPatGeneral)) return general32(r.ex(i), b, L);
  if (kb & kPatSquare) return kind2(b);
  return kind1(b, vx);
}

// t <- t * (q - r)^e, with the product taken inside each kind's branch so the
// branches share no term registers (this file is compiled without
// SimplifyCFG's common-code sinking, which would merge the products behind
// register copies).
template <bool kGeneral, class Acc>
__device__ __forceinline__ void apply32(q32& t, const F4& rt, int kb, const Acc& r, int i, const q32& q, float& L,
                                        float& vx) {
  const q32 b = qsub32(q, rt);
  if (kGeneral && (kb & kPatGeneral)) {
    t = qmul32(t, general32(r.ex(i), b, L));
  } else if (kb & kPatSquare) {
    t = qmul32(t, kind2(b));
  } else {
    t = qmul32(t, kind1(b, vx));
  }
}

template <class QP>
__device__ __forceinline__ F4 load_root(QP p) {
  return F4{p->w, p->x, p->y, p->z};
}

// Roots through pointers into a root block (generic, or the constant address
// space where a kernel wants scalar loads of a pointer it read from memory).
template <class QP, class PP, class EP>
struct RootAcc {
  QP q;
  PP p;
  EP e;
  __device__ __forceinline__ F4 root(int i) const { return load_root(q + i); }
  __device__ __forceinline__ int pat(int i) const { return p[i]; }
  __device__ __forceinline__ float ex(int i) const { return e[i]; }
};
template <class QP, class PP, class EP>
__device__ __forceinline__ RootAcc<QP, PP, EP> root_acc(QP q, PP p, EP e, int off) {
  return RootAcc<QP, PP, EP>{q + off, p + off, e + off};
}

template <int kRN, bool kGeneral, class Acc>
__device__ __forceinline__ q32 poly32(const Acc& r, int n, const q32& q, int& E, float& L, float& hi, float& lo,
                                      bool& bad, float& vx) {
  q32 t = term32<kGeneral>(r.root(0), r.pat(0), r, 0, q, L, vx);
  int mn = 0;
  int ehi = -(1 << 30), elo = 1 << 30;
  // roots come in sub-groups of up to 4: one scalar load of four roots and
  // one pattern word (the kinds of roots i..i+3) each
  constexpr int kSub = kRN < 4 ? kRN : 4;
  auto group = [&](int i) {
#pragma unroll
    for (int s0 = 0; s0 < kRN; s0 += kSub) {
      F4 g[kSub];
#pragma unroll
      for (int j = 0; j < kSub; j++) g[j] = r.root(i + s0 + j);
      const int pat = r.pat(i + s0);
#pragma unroll
      for (int j = 0; j < kSub; j++) apply32<kGeneral>(t, g[j], pat >> (2 * j), r, i + s0 + j, q, L, vx);
    }
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
  };
  int i = 1;
  for (; i + kRN - 1 < n; i += kRN) group(i);
  for (; i < n; i++) apply32<kGeneral>(t, r.root(i), r.pat(i), r, i, q, L, vx);
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
// Every 8 terms (integer exponents only): the first group holds the first
// term and 8 more, 9 * (15 - 2) + 1 < 127 (term_bits carries a 2-bit margin).
// A running product that underflowed inside a group must end the group below
// 2^-60 (the check that sends it to FP64): a product of at most 9 terms
// starting at >= 2^-1 falls under 2^-126 only if some term is below 2^-14,
// that is q is within 1 of a root, and then every term is at most 2^near_bits,
// so the group ends below 2^(-126 + 8 near_bits) <= 2^-62 when near_bits <= 8.
constexpr double kRN8MaxTermBits = 15.0;
constexpr double kRN8MaxNearBits = 8.0;
// FP32 variant for an integer-exponent map: terms between renormalisations.
// 8-term groups only pay when a polynomial has at least one full group after
// its first term.
__host__ __device__ inline int rn_of(const MapArgs& a) {
  if (a.term_bits <= kRN8MaxTermBits && a.near_bits <= kRN8MaxNearBits && (a.nt > 8 || a.nb > 8)) return 8;
  return a.term_bits <= kRN4MaxTermBits ? 4 : 2;
}

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
// edge, or outside it, goes to FP64. With QJ_MAP_EXTENDED_RANGE there is no
// range edge (the products' exponents are integers): only an iterate that
// continues below 2^-100, where FP32 would lose its precision, goes to FP64
// (the extended-range FP64 evaluator).
// top / bot: root accessors of the two polynomials (a.nt and a.nb roots).
template <int kRN, bool kGeneral, class AccT, class AccB>
__device__ __forceinline__ Res32 eval32(const MapArgs& a, const AccT& top, const AccB& bot, const Consts32& k,
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
  const bool ext = (a.flags & QJ_MAP_EXTENDED_RANGE) != 0;
  q32 q{(float)q0.w, (float)q0.x, (float)q0.y, (float)q0.z};
  float log2mag = 0.0f;
  float vx = __builtin_inff();
  bool amb = false, bad = false, zero = false;
  int it = 0;
  // FP64 running products between two checkpoints grow by at most (kRN - 1)
  // terms; with 8-term groups a product may also dip up to 60 + 7 near_bits
  // (<= 116) bits below a checkpoint before the group ends, so the low side
  // keeps a further 64 bits from FP64's range
  const float lim = 1020.0f - (float)(kRN - 1) * k.term_bits;
  const float lim_lo = lim - (kRN > 4 ? 64.0f : 0.0f);
  while (true) {
    int ET = 0, EB = 0;
    float L = 0.0f, hiT = -1e30f, loT = 1e30f;
    const q32 T = poly32<kRN, kGeneral>(top, a.nt, q, ET, L, hiT, loT, bad, vx);
    bad |= !ext && !(hiT < lim && loT > -lim_lo);
    float l2 = k.log2c;
    if (a.nb > 0) {
      float LB = 0.0f, hiB = -1e30f, loB = 1e30f;
      const q32 B = poly32<kRN, kGeneral>(bot, a.nb, q, EB, LB, hiB, loB, bad, vx);
      bad |= !ext && !(hiB < lim && loB > -lim_lo);
      const float nb2 = B.w * B.w + B.x * B.x + B.y * B.y + B.z * B.z;
      const float lB = 0.5f * __log2f(nb2) + (float)EB + LB;
      const q32 D = qmul32(T, q32{B.w, -B.x, -B.y, -B.z});  // T conj(B) = T B^-1 |B|^2
      const float nd2 = D.w * D.w + D.x * D.x + D.y * D.y + D.z * D.z;
      l2 += 0.5f * __log2f(nd2) - __log2f(nb2) + (float)(ET - EB) + (L - LB);
      bad |= !(nb2 > 0.0f) || !(nd2 > 0.0f);
      if (!ext && lB > 512.5f) {
        zero = true;  // FP64: |B|^2 = inf, 1/inf = 0, q = T * 0 = 0
      } else if (!ext && (lB > 511.5f || lB < -510.0f)) {
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
      log2mag = -__builtin_inff();
      break;
    }
    log2mag = l2;
    if (ext) bad |= !isfinite(l2);
    else bad |= !isfinite(l2) || l2 > 511.0f || l2 < -530.0f;  // |q|^2 over/underflow in FP64
    amb |= fabsf(l2 - k.log2esc) < k.margin;
    const bool bh = black_hole && l2 < k.log2bh;
    if (black_hole) amb |= fabsf(l2 - k.log2bh) < k.margin;
    if (bad || l2 >= k.log2esc || bh || it >= a.max_iter) break;
    if (ext && l2 < -100.0f) {
      bad = true;  // the next iteration would start below FP32's precise range
      break;
    }
  }
  res.F = zero ? -__builtin_inff() : log2mag * 0.69314718f;
  res.need64 = bad || amb || !(vx > 0.0f) || !(fabsf(res.F) >= k.band);
  return res;
}

// ---------------------------------------------------------------- field tiles

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

// FP32 bulk of one field tile per block (the body of k_grid32).
template <int kRN, bool kGeneral, class AccT, class AccB>
__device__ __forceinline__ void grid32_body(const MapArgs& a, const Consts32& k, const AccT& top, const AccB& bot,
                                            double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                            uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount,
                                            uint32_t bcap) {
  int x, y, z;
  const bool valid = tile_voxel(a, x, y, z);
  int iters = 0;
  bool need = false;
  const int64_t idx = x + (int64_t)y * a.res[0] + (int64_t)z * a.res[0] * a.res[1];
  if (valid) {
    double c[3];
    cell_center(a, x, y, z + a.zoff, c);
    const Res32 r = eval32<kRN, kGeneral>(a, top, bot, k, q64{c[0], c[1], c[2], a.slice});
    if (r.written) {
      if (!r.need64) out[idx] = (double)r.F;
      need = r.need64;
    }
    iters = r.iters;
  }
  block_count(valid ? 1 : 0, cnt, CNT_VOXELS);
  block_count(iters, cnt, CNT_ITERS);
  if (!block_append(need, (uint32_t)idx, blist, bcount, bcap) && need) ((unsigned long long*)out)[idx] = kBandMark;
}

}  // namespace qj
