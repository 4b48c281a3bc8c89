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
//   k_grid64        FP64, reference expression order (the "plain FP64" method).
//   k_grid32        FP32 bulk with exact power-of-two range scaling (device
//                   code in qj_fp32.h); marks the band |F| < band, non-finite
//                   results and near-threshold escape / black-hole decisions
//                   into a compacted list (wave ballot + one atomic per block).
//   k_points32      the same for a list of points.
//   k_list64        FP64 re-evaluation of a compacted list (band or points);
//                   k_list64<.., true> runs eval64x, FP64 with an explicit
//                   binary exponent (extended range) for cells that
//                   overflowed: the reference's long double re-pass.
//   k_batch32, k_batch64_list, k_batch64
//                   the request loop's work-queue kernels over chunks of
//                   many requests (FP32 bulk per exponent variant, FP64 band
//                   lists, plain FP64).
#include <float.h>

#include "qj_geom.h"
#include "qj_internal.h"
#include "qj_quat.h"
#include "qj_eval.h"
#include "qj_fp32.h"

namespace qj {


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

// ---------------------------------------------------------------- FP64 extended range (AP re-pass)

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

// ---------------------------------------------------------------- kernels


// kLagrange: the map may be in Lagrange order (QJ_MAP_LAGRANGE); the
// factored-order instance carries no Lagrange code and needs fewer registers.
// kPow: kPolar is the reference method (QJ_FP64); kExactOnly / kExactInt take
// exponents 1 and 2 as exact products (QJ_FP64_EXACT). kExt: the
// extended-range evaluator (QJ_MAP_EXTENDED_RANGE).
template <bool kLagrange, int kPow = kPolar, bool kExt = false>
__global__ __launch_bounds__(kBlock) void k_grid64(MapArgs a, const RootRec64* __restrict__ roots,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   uint32_t* __restrict__ alist, uint32_t* __restrict__ acount,
                                                   uint32_t acap) {
  int x, y, z;
  const bool valid = tile_voxel(a, x, y, z);
  int iters = 0;
  bool ovf = false, nonfinite = false;
  const int64_t idx = x + (int64_t)y * a.res[0] + (int64_t)z * a.res[0] * a.res[1];
  if (valid) {
    double c[3];
    cell_center(a, x, y, z + a.zoff, c);
    const q64 q0{c[0], c[1], c[2], a.slice};
    const Res64 r = kExt ? eval64e<kPow>(a, roots, q0) : eval64<kPow, const RootRec64*, kLagrange>(a, roots, q0);
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
    block_append(ovf, (uint32_t)idx, alist, acount, acap);
    block_count(ovf ? 1 : 0, cnt, CNT_ADAPTIVE);
  }
}

template <int kRN, bool kGeneral>
__global__ __launch_bounds__(kBlock) void k_grid32(MapArgs a, Consts32 k, const F4* __restrict__ roots,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount,
                                                   uint32_t bcap) {
  const Roots32 R = roots32_view(roots, a.nt + a.nb);
  grid32_body<kRN, kGeneral>(a, k, root_acc(R.q, R.pat, R.e, 0), root_acc(R.q, R.pat, R.e, a.nt), out, cnt,
                                    blist, bcount, bcap);
}

template <class Src, int kRN, bool kGeneral>
__global__ __launch_bounds__(kBlock) void k_points32(MapArgs a, Consts32 k, const F4* __restrict__ roots,
                                                     const double* __restrict__ pts, int64_t n,
                                                     double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                     uint32_t* __restrict__ blist, uint32_t* __restrict__ bcount,
                                                     uint32_t bcap) {
  const Roots32 R = roots32_view(roots, a.nt + a.nb);
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  int iters = 0;
  bool need = false;
  if (valid) {
    const Res32 r = eval32<kRN, kGeneral>(a, root_acc(R.q, R.pat, R.e, 0), root_acc(R.q, R.pat, R.e, a.nt), k,
                                                Src::start(a, pts, i));
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
  if (!block_append(need, (uint32_t)i, blist, bcount, bcap) && need) ((unsigned long long*)out)[i] = kBandMark;
}

// FP64 over a compacted list (count read on the device, at most cap entries
// held), or over all n indices when list == nullptr. kPow: the PowMode
// (kExactInt / kExactOnly for the FP32 path's band and range re-evaluation).
template <class Src, bool kExtended, int kPow = kPolar>
__global__ __launch_bounds__(kBlock) void k_list64(MapArgs a, const RootRec64* __restrict__ roots,
                                                   const double* __restrict__ pts,
                                                   const uint32_t* __restrict__ list,
                                                   const uint32_t* __restrict__ count, int64_t cap, int64_t n_all,
                                                   double* __restrict__ out, unsigned long long* __restrict__ cnt,
                                                   int slot_vox, int slot_iter, uint32_t* __restrict__ alist,
                                                   uint32_t* __restrict__ acount, uint32_t acap) {
  const int64_t n = list ? ((int64_t)*count < cap ? (int64_t)*count : cap) : n_all;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  unsigned long long vox = 0, iters = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n; base += stride) {
    const int64_t i = base + threadIdx.x;
    bool ovf = false;
    uint32_t idx = 0;
    if (i < n) {
      idx = list ? list[i] : (uint32_t)i;
      const q64 q0 = Src::start(a, pts, idx);
      const Res64 r = kExtended ? eval64x(a, roots, q0) : eval64_any<kPow>(a, roots, q0);
      if (r.written) out[idx] = r.F;
      else if (pts) out[idx] = __builtin_nan("");
      vox++;
      iters += r.iters;
      ovf = r.overflow;
    }
    if (!kExtended && (a.flags & QJ_MAP_ADAPTIVE)) block_append(ovf, idx, alist, acount, acap);
  }
  block_count(vox, cnt, slot_vox);
  block_count(iters, cnt, slot_iter);
}

// The cells (or points) that the FP32 pass found in need of FP64 after the
// band list had filled up (they hold kBandMark), evaluated as k_list64
// evaluates the listed ones. Returns at once when the list held them all.
template <class Src, int kPow>
__global__ __launch_bounds__(kBlock) void k_band_sweep(MapArgs a, const RootRec64* __restrict__ roots,
                                                       const double* __restrict__ pts,
                                                       const uint32_t* __restrict__ count, int64_t cap,
                                                       int64_t n_all, double* __restrict__ out,
                                                       unsigned long long* __restrict__ cnt,
                                                       uint32_t* __restrict__ alist, uint32_t* __restrict__ acount,
                                                       uint32_t acap) {
  if ((int64_t)*count <= cap) return;  // the same for every thread of the grid
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  unsigned long long vox = 0, iters = 0;
  for (int64_t base = (int64_t)blockIdx.x * kBlock; base < n_all; base += stride) {
    const int64_t i = base + threadIdx.x;
    bool ovf = false;
    if (i < n_all && ((const unsigned long long*)out)[i] == kBandMark) {
      const q64 q0 = Src::start(a, pts, (uint32_t)i);
      const Res64 r = eval64_any<kPow>(a, roots, q0);
      if (r.written) out[i] = r.F;
      else if (pts) out[i] = __builtin_nan("");
      vox++;
      iters += r.iters;
      ovf = r.overflow;
    }
    if (a.flags & QJ_MAP_ADAPTIVE) block_append(ovf, (uint32_t)i, alist, acount, acap);
  }
  block_count(vox, cnt, CNT_BAND_VOXELS);
  block_count(iters, cnt, CNT_BAND_ITERS);
}

// ---------------------------------------------------------------- sampled FP64 cross-check

// Block maximum of a non-negative double into counter slot `slot` (the bits
// of non-negative doubles order like unsigned integers).
template <int kSpread = kCounterSpread>
__device__ __forceinline__ void block_max(double v, unsigned long long* cnt, int slot) {
  __shared__ unsigned long long s_part[kBlock / 64];
  unsigned long long b = __builtin_bit_cast(unsigned long long, v > 0.0 ? v : 0.0);
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(b, o);
    b = t > b ? t : b;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) s_part[wave] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long m = 0;
    for (int i = 0; i < kBlock / 64; i++) m = s_part[i] > m ? s_part[i] : m;
    if (m) atomicMax(&cnt[slot * kSpread + (blockIdx.x % kSpread)], m);
  }
  __syncthreads();
}

// The error statement of a finished field: at `samples` evenly spaced cells
// the reference method is evaluated again in FP64 (polar powers, reference
// expression order; the extended-range re-pass where QJ_MAP_ADAPTIVE flags an
// overflow) and compared with the value written: counts of checked cells,
// sign disagreements of (float)F > 0 (the mesher's predicate) and class
// disagreements (finite / NaN / +-inf), and the largest absolute and relative
// (to max(1, |F64|)) differences over cells finite in both.
// kSpread: copies per counter slot (a workspace's counters), or 1 for the
// request loop's per-result block of CNT_N counters.
template <int kSpread = kCounterSpread>
__global__ __launch_bounds__(kBlock) void k_check64(MapArgs a, const RootRec64* __restrict__ roots,
                                                    const double* __restrict__ field, int64_t nvox, int64_t samples,
                                                    unsigned long long* __restrict__ cnt) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  unsigned long long checked = 0, flips = 0, cls = 0;
  double dabs = 0.0, drel = 0.0;
  if (i < samples) {
    const int64_t idx = ((2 * i + 1) * nvox) / (2 * samples);
    const q64 q0 = GridSource::start(a, nullptr, idx);
    Res64 r = (a.flags & QJ_MAP_EXTENDED_RANGE) ? eval64e<kPolar>(a, roots, q0)
                                                : eval64<kPolar, const RootRec64*, true>(a, roots, q0);
    if ((a.flags & QJ_MAP_ADAPTIVE) && r.overflow) r = eval64x(a, roots, q0);
    if (r.written) {
      const double F = field[idx];
      checked = 1;
      const bool fin = isfinite(F), fin64 = isfinite(r.F);
      if (fin != fin64 || isnan(F) != isnan(r.F) || (!fin && !isnan(F) && F != r.F)) {
        cls = 1;
      } else if (fin) {
        dabs = fabs(F - r.F);
        drel = dabs / fmax(1.0, fabs(r.F));
      }
      flips = (((float)F > 0.0f) != ((float)r.F > 0.0f)) ? 1 : 0;
    }
  }
  block_count<kSpread>(checked, cnt, CNT_CHECKED);
  block_count<kSpread>(flips, cnt, CNT_CHECK_FLIPS);
  block_count<kSpread>(cls, cnt, CNT_CHECK_CLASS);
  block_max<kSpread>(dabs, cnt, CNT_CHECK_MAXABS);
  block_max<kSpread>(drel, cnt, CNT_CHECK_MAXREL);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

qj_status launch_check_one(const MapArgs& a, const RootRec64* roots, const double* field, int64_t samples,
                           unsigned long long* d_out, hipStream_t s) {
  const int64_t nvox = (int64_t)a.res[0] * a.res[1] * a.res[2];
  const int64_t m = samples < nvox ? samples : nvox;
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, s, d_out, (int)CNT_N);
  QJ_HIP(hipGetLastError());
  if (m <= 0) return QJ_OK;
  hipLaunchKernelGGL(k_check64<1>, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a, roots, field,
                     nvox, m, d_out);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
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
// the front), [1] band count of general chunks (list from the back), [2..5]
// FP32 queues per variant, [6] FP64 queue.
enum { BC_BAND_INT = 0, BC_BAND_GEN = 1, BC_Q32 = 2, BC_Q64 = 6, BC_N = 8 };

// Descriptors are grouped by FP32 variant: chunks [0, g1) have general
// exponents, [g1, g2) integer exponents renormalised every 4 terms, [g2, g3)
// every 2 and [g3, total) every 8.
template <int kVar>
__global__ __launch_bounds__(kBlock) void k_batch32(const ChunkDesc* __restrict__ descs, int blocks_per_chunk,
                                                    int chunk_voxels, uint32_t* __restrict__ blist, int64_t list_cap,
                                                    uint32_t* __restrict__ counters) {
  constexpr int kRN = kVar == 1 ? 4 : (kVar == 3 ? 8 : 2);
  constexpr bool kGen = kVar == 0;
  const ChunkDesc& d0 = descs[0];
  const int c0 = kVar == 0 ? 0 : (kVar == 1 ? d0.g1 : (kVar == 2 ? d0.g2 : d0.g3));
  const int c1 = kVar == 0 ? d0.g1 : (kVar == 1 ? d0.g2 : (kVar == 2 ? d0.g3 : d0.total));
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
      if (__builtin_isinf(k.band)) {
        // a band of infinite width (a result escalated to FP64): every cell goes
        // to the FP64 pass, so the FP32 evaluation is skipped; cells that eval32
        // leaves unwritten (the skipped origin) stay out of the list
        const q64 q0 = chunk_start(a, idx);
        need = !((a.flags & QJ_MAP_SKIP_ORIGIN) && qmag(q0) < 1e-8);
      } else {
        const Roots32K R = roots32_const(d.r32, a.nt + a.nb);
        const Res32 r = eval32<kRN, kGen>(a, root_acc(R.q, R.pat, R.e, 0), root_acc(R.q, R.pat, R.e, a.nt), k,
                                               chunk_start(a, idx));
        if (r.written) {
          if (!r.need64) d.out[idx] = (double)r.F;
          need = r.need64;
        }
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
      block_append(need, e, blist, counters + BC_BAND_INT, (uint32_t)list_cap);
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
    // the list interleaves the chunks that were in flight together, but a
    // wave's entries nearly always come from one request: chunks of a request
    // share its map arguments, roots and output, so those are then
    // wave-uniform and the root records come in scalar loads; only the
    // chunk's first voxel is per lane
    const int cw = __builtin_amdgcn_readfirstlane(c);
    const ChunkDesc& dw = descs[cw];
    const ChunkDesc& dl = descs[c];
    if (__ballot(dl.r64 != dw.r64) == 0ull) {
      const int64_t idx = dl.start + local;
      typedef const __attribute__((address_space(4))) RootRec64 cR64;
      const Res64 r = eval64_any<kPow>(dw.a, (cR64*)dw.r64, chunk_start(dw.a, idx));
      if (r.written) dw.out[idx] = r.F;
    } else {
      const ChunkDesc& d = descs[c];
      const int64_t idx = d.start + local;
      const Res64 r = eval64_any<kPow>(d.a, d.r64, chunk_start(d.a, idx));
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
      const Res64 r = eval64_any(d.a, d.r64, chunk_start(d.a, idx));
      if (r.written) d.out[idx] = r.F;
    }
  }
}

// The launch's descriptors from pinned host memory (coherent, read over the
// bus) into device memory: the first node of every request-loop graph. Only
// the descriptors in use (the first descriptor's total) are copied.
__global__ __launch_bounds__(kBlock) void k_desc_fetch(const ChunkDesc* __restrict__ h_desc,
                                                       ChunkDesc* __restrict__ d_desc, int max_chunks) {
  static_assert(sizeof(ChunkDesc) % 8 == 0, "descriptors copy in 8-byte words");
  const int total = h_desc[0].total;
  const int n = total < max_chunks ? (total > 1 ? total : 1) : max_chunks;
  const int64_t words = (int64_t)n * (int64_t)(sizeof(ChunkDesc) / 8);
  const unsigned long long* src = (const unsigned long long*)h_desc;
  unsigned long long* dst = (unsigned long long*)d_desc;
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < words; i += (int64_t)gridDim.x * kBlock)
    dst[i] = src[i];
}

qj_status launch_desc_fetch(const ChunkDesc* h_desc, ChunkDesc* d_desc, int max_chunks, hipStream_t s) {
  const int64_t words = (int64_t)max_chunks * (int64_t)(sizeof(ChunkDesc) / 8);
  const int64_t blocks = (words + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_desc_fetch, dim3((unsigned)(blocks < 1 ? 1 : blocks)), dim3(kBlock), 0, s, h_desc, d_desc,
                     max_chunks);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

static int resident_blocks(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 2048;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) return 2048;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu <= 0)
    return 2 * cus;
  return per_cu * cus;
}

int batch_variant(const MapArgs& a) {
  if (a.gen) return 0;
  const int rn = rn_of(a);
  return rn == 8 ? 3 : (rn == 4 ? 1 : 2);
}

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
  static const int g3 = resident_blocks((const void*)k_batch32<3>);
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
  if (variants & 8u) {
    hipLaunchKernelGGL(k_batch32<3>, dim3(g3), dim3(kBlock), 0, s, d_desc, bpc, chunk_voxels, d_band, cap, d_counters);
    QJ_HIP(hipGetLastError());
  }
  if (variants & 1u) {
    hipLaunchKernelGGL(k_batch32<0>, dim3(g0), dim3(kBlock), 0, s, d_desc, bpc, chunk_voxels, d_band, cap, d_counters);
    QJ_HIP(hipGetLastError());
  }
  if (variants & 14u) {  // band voxels of the integer-exponent variants
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

// Device buffers of one field evaluation: the band list (list_cap entries;
// what does not fit is marked in the output and swept), the adaptive list
// (alist_cap entries, as many as cells when QJ_MAP_ADAPTIVE is set), their two
// counts, and the statistics counters. ws (may be null) only carries the
// optional kernel timing.
struct FieldBufs {
  uint32_t* list;
  uint32_t* alist;
  int64_t list_cap;
  int64_t alist_cap;
  uint32_t* list_count;
  unsigned long long* counters;
  qj_workspace* ws;
};

static FieldBufs ws_bufs(qj_workspace* ws) {
  return FieldBufs{ws->d_list, ws->d_alist, ws->list_cap, ws->alist_cap, ws->d_list_count, ws->d_counters, ws};
}

// The band sweep after the band list's FP64 pass (exits at once unless the
// FP32 pass overflowed the list).
template <class Src>
static qj_status launch_sweep(const FieldBufs& fb, const MapArgs& a, const qj_rational* f, const double* pts,
                              int64_t n, double* d_out, hipStream_t s) {
  const dim3 g(list_grid(n / 16)), b(kBlock);
  if (a.gen)
    hipLaunchKernelGGL((k_band_sweep<Src, kExactInt>), g, b, 0, s, a, f->d64, pts, fb.list_count, fb.list_cap, n,
                       d_out, fb.counters, fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
  else
    hipLaunchKernelGGL((k_band_sweep<Src, kExactOnly>), g, b, 0, s, a, f->d64, pts, fb.list_count, fb.list_cap, n,
                       d_out, fb.counters, fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

static void mark(const FieldBufs& b, hipStream_t s) {
  if (b.ws) timing_mark(b.ws, s);
}

static qj_status run_adaptive(const FieldBufs& fb, const MapArgs& a, const qj_rational* f, const double* pts,
                              int64_t n, double* d_out, hipStream_t s) {
  if (!(a.flags & QJ_MAP_ADAPTIVE)) return QJ_OK;
  if (pts)
    hipLaunchKernelGGL((k_list64<PointSource, true>), dim3(list_grid(n / 64 + 1)), dim3(kBlock), 0, s, a, f->d64,
                       pts, fb.alist, fb.list_count + 1, fb.alist_cap, (int64_t)0, d_out, fb.counters,
                       (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u);
  else
    hipLaunchKernelGGL((k_list64<GridSource, true>), dim3(list_grid(n / 64 + 1)), dim3(kBlock), 0, s, a, f->d64,
                       pts, fb.alist, fb.list_count + 1, fb.alist_cap, (int64_t)0, d_out, fb.counters,
                       (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, (uint32_t*)nullptr, (uint32_t*)nullptr, 0u);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

// One field (or one z-slab of it: a.res[2] slices from global slice a.zoff,
// written from d_out on); with QJ_MAP_ADAPTIVE the adaptive list holds at
// least a.res[0] a.res[1] a.res[2] entries.
static qj_status launch_field_bufs(const FieldBufs& fb, const MapArgs& a, const qj_rational* f, qj_precision prec,
                                   double band, double* d_out, hipStream_t s) {
  const int64_t nvox = (int64_t)a.res[0] * a.res[1] * a.res[2];
  // list counters cleared by a kernel, not a memset node, so the call can be captured in a graph
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, fb.list_count, 2);
  QJ_HIP(hipGetLastError());
  const int64_t ntx = (a.res[0] + kTileX - 1) / kTileX, nty = (a.res[1] + kTileY - 1) / kTileY,
                ntz = (a.res[2] + kTileZ - 1) / kTileZ;
  const int64_t nblocks = ntx * nty * ntz;
  if (nblocks > INT32_MAX) return QJ_ERR_CAPACITY;
  mark(fb, s);
  const bool ext = (a.flags & QJ_MAP_EXTENDED_RANGE) != 0;
  const dim3 g64((unsigned)nblocks), b64(kBlock);
  if (prec == QJ_FP64_EXACT) {
    if (a.gen && ext)
      hipLaunchKernelGGL((k_grid64<false, kExactInt, true>), g64, b64, 0, s, a, f->d64, d_out, fb.counters, fb.alist,
                         fb.list_count + 1, (uint32_t)fb.alist_cap);
    else if (a.gen)
      hipLaunchKernelGGL((k_grid64<false, kExactInt>), g64, b64, 0, s, a, f->d64, d_out, fb.counters, fb.alist,
                         fb.list_count + 1, (uint32_t)fb.alist_cap);
    else if (ext)
      hipLaunchKernelGGL((k_grid64<false, kExactOnly, true>), g64, b64, 0, s, a, f->d64, d_out, fb.counters,
                         fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
    else
      hipLaunchKernelGGL((k_grid64<false, kExactOnly>), g64, b64, 0, s, a, f->d64, d_out, fb.counters, fb.alist,
                         fb.list_count + 1, (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
    mark(fb, s);
    mark(fb, s);
  } else if (prec == QJ_FP64 && ext) {
    hipLaunchKernelGGL((k_grid64<false, kPolar, true>), g64, b64, 0, s, a, f->d64, d_out, fb.counters, fb.alist,
                       fb.list_count + 1, (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
    mark(fb, s);
    mark(fb, s);
  } else if (prec == QJ_FP64) {
    if (a.flags & QJ_MAP_LAGRANGE)
      hipLaunchKernelGGL(k_grid64<true>, dim3((unsigned)nblocks), dim3(kBlock), 0, s, a, f->d64, d_out,
                         fb.counters, fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
    else
      hipLaunchKernelGGL(k_grid64<false>, dim3((unsigned)nblocks), dim3(kBlock), 0, s, a, f->d64, d_out,
                         fb.counters, fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
    mark(fb, s);
    mark(fb, s);
  } else {
    const Consts32 k = consts32(a, band);
    const dim3 g((unsigned)nblocks), b(kBlock);
    if (a.gen)
      hipLaunchKernelGGL((k_grid32<2, true>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, fb.counters, fb.list,
                         fb.list_count, (uint32_t)fb.list_cap);
    else if (rn_of(a) == 8)
      hipLaunchKernelGGL((k_grid32<8, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, fb.counters, fb.list,
                         fb.list_count, (uint32_t)fb.list_cap);
    else if (rn_of(a) == 4)
      hipLaunchKernelGGL((k_grid32<4, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, fb.counters, fb.list,
                         fb.list_count, (uint32_t)fb.list_cap);
    else
      hipLaunchKernelGGL((k_grid32<2, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_out, fb.counters, fb.list,
                         fb.list_count, (uint32_t)fb.list_cap);
    QJ_HIP(hipGetLastError());
    mark(fb, s);
    mark(fb, s);
    if (a.gen)
      hipLaunchKernelGGL((k_list64<GridSource, false, kExactInt>), dim3(list_grid(nvox / 16)), dim3(kBlock), 0, s, a,
                         f->d64, (const double*)nullptr, fb.list, fb.list_count, fb.list_cap, (int64_t)0, d_out,
                         fb.counters, (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    else
      hipLaunchKernelGGL((k_list64<GridSource, false, kExactOnly>), dim3(list_grid(nvox / 16)), dim3(kBlock), 0, s,
                         a, f->d64, (const double*)nullptr, fb.list, fb.list_count, fb.list_cap, (int64_t)0, d_out,
                         fb.counters, (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
    qj_status st = launch_sweep<GridSource>(fb, a, f, nullptr, nvox, d_out, s);
    if (st) return st;
  }
  mark(fb, s);
  return run_adaptive(fb, a, f, nullptr, nvox, d_out, s);
}

qj_status launch_field(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec, double band,
                       double* d_out, hipStream_t s) {
  const int64_t nvox = (int64_t)a.res[0] * a.res[1] * a.res[2];
  if (nvox <= 0) return QJ_OK;
  if (nvox > (int64_t)UINT32_MAX) {
    set_error("grid of %lld cells exceeds 2^32; evaluate it in tiles", (long long)nvox);
    return QJ_ERR_CAPACITY;
  }
  qj_status st = ensure_list(ws, nvox, (a.flags & QJ_MAP_ADAPTIVE) != 0, s);
  if (st) return st;
  if ((st = launch_field_bufs(ws_bufs(ws), a, f, prec, band, d_out, s))) return st;
  if (ws->check_samples > 0) {
    const int64_t m = ws->check_samples < nvox ? ws->check_samples : nvox;
    hipLaunchKernelGGL(k_check64<>, dim3((unsigned)((m + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, a, f->d64,
                       (const double*)d_out, nvox, m, ws->d_counters);
    QJ_HIP(hipGetLastError());
  }
  return QJ_OK;
}

qj_status launch_points(qj_workspace* ws, const MapArgs& a, const qj_rational* f, qj_precision prec, double band,
                        const double* d_pts, int64_t n, double* d_out, hipStream_t s) {
  if (n <= 0) return QJ_OK;
  if (n > (int64_t)UINT32_MAX) return QJ_ERR_CAPACITY;
  qj_status st = ensure_list(ws, n, (a.flags & QJ_MAP_ADAPTIVE) != 0, s);
  if (st) return st;
  const FieldBufs fb = ws_bufs(ws);
  // list counters cleared by a kernel, not a memset node, so the call can be captured in a graph
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, s, fb.list_count, 2);
  QJ_HIP(hipGetLastError());
  if (prec == QJ_FP64) {
    hipLaunchKernelGGL((k_list64<PointSource, false>), dim3(list_grid(n)), dim3(kBlock), 0, s, a, f->d64, d_pts,
                       (const uint32_t*)nullptr, (const uint32_t*)nullptr, (int64_t)0, n, d_out, fb.counters,
                       (int)CNT_VOXELS, (int)CNT_ITERS, fb.alist, fb.list_count + 1, (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
  } else if (prec == QJ_FP64_EXACT) {
    if (a.gen)
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactInt>), dim3(list_grid(n)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, (const uint32_t*)nullptr, (const uint32_t*)nullptr, (int64_t)0, n, d_out,
                         fb.counters, (int)CNT_VOXELS, (int)CNT_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    else
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactOnly>), dim3(list_grid(n)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, (const uint32_t*)nullptr, (const uint32_t*)nullptr, (int64_t)0, n, d_out,
                         fb.counters, (int)CNT_VOXELS, (int)CNT_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
  } else {
    const int64_t nb = (n + kBlock - 1) / kBlock;
    const Consts32 k = consts32(a, band);
    const dim3 g((unsigned)nb), b(kBlock);
    if (a.gen)
      hipLaunchKernelGGL((k_points32<PointSource, 2, true>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         fb.counters, fb.list, fb.list_count, (uint32_t)fb.list_cap);
    else if (rn_of(a) == 8)
      hipLaunchKernelGGL((k_points32<PointSource, 8, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         fb.counters, fb.list, fb.list_count, (uint32_t)fb.list_cap);
    else if (rn_of(a) == 4)
      hipLaunchKernelGGL((k_points32<PointSource, 4, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         fb.counters, fb.list, fb.list_count, (uint32_t)fb.list_cap);
    else
      hipLaunchKernelGGL((k_points32<PointSource, 2, false>), g, b, 0, s, a, k, (const F4*)f->d32, d_pts, n, d_out,
                         fb.counters, fb.list, fb.list_count, (uint32_t)fb.list_cap);
    QJ_HIP(hipGetLastError());
    if (a.gen)
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactInt>), dim3(list_grid(n / 16)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, fb.list, fb.list_count, fb.list_cap, (int64_t)0, d_out, fb.counters,
                         (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    else
      hipLaunchKernelGGL((k_list64<PointSource, false, kExactOnly>), dim3(list_grid(n / 16)), dim3(kBlock), 0, s, a,
                         f->d64, d_pts, fb.list, fb.list_count, fb.list_cap, (int64_t)0, d_out, fb.counters,
                         (int)CNT_BAND_VOXELS, (int)CNT_BAND_ITERS, fb.alist, fb.list_count + 1,
                         (uint32_t)fb.alist_cap);
    QJ_HIP(hipGetLastError());
    if ((st = launch_sweep<PointSource>(fb, a, f, d_pts, n, d_out, s))) return st;
  }
  return run_adaptive(fb, a, f, d_pts, n, d_out, s);
}

}  // namespace qj

// ---------------------------------------------------------------- allocation-free entry point

namespace qj {

// Scratch of qj_field_eval_noalloc, a file-scope device array (each GPU holds
// its own copy until the process ends): for one z-slab of at most
// kStaticSlabVoxels cells, the band list and the adaptive list, their two
// counts, and the statistics counters the kernels add to. 32 MiB + 4.3 KB.
constexpr int64_t kStaticSlabVoxels = int64_t(1) << 22;
constexpr size_t kStaticListBytes = sizeof(uint32_t) * (size_t)kStaticSlabVoxels;
constexpr size_t kStaticCountOff = 2 * kStaticListBytes;
constexpr size_t kStaticCounterOff = kStaticCountOff + 256;
constexpr size_t kStaticBytes = kStaticCounterOff + sizeof(unsigned long long) * CNT_N * kCounterSpread;
static __device__ __align__(256) unsigned char scratch[kStaticBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch[kMaxDevices];  // device address of `scratch` per GPU

// Looks up the address of this GPU's copy of `scratch`; `device` must be the
// current device. Host code uses only that address, never the array's name.
qj_status scratch_setup(int device, unsigned char** out) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d beyond the %d the allocation-free entry point supports", device, kMaxDevices);
    return QJ_ERR_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch[device] = (unsigned char*)p;
  }
  if (out) *out = g_scratch[device];
  return QJ_OK;
}

}  // namespace qj

extern "C" qj_status qj_field_eval_noalloc(const qj_grid* grid, const qj_rational* f, const qj_map_params* params,
                                           qj_precision precision, double band, double* d_field, void* stream) {
  using namespace qj;
  if (!d_field) {
    set_error("qj_field_eval_noalloc: null output");
    return QJ_ERR_INVALID;
  }
  if (precision != QJ_FP64 && precision != QJ_FP32_BAND && precision != QJ_FP64_EXACT) {
    set_error("qj_field_eval_noalloc: unknown precision %d", (int)precision);
    return QJ_ERR_INVALID;
  }
  if (params && (params->flags & QJ_MAP_LAGRANGE) && precision != QJ_FP64) {
    set_error("qj_field_eval_noalloc: QJ_MAP_LAGRANGE needs QJ_FP64");
    return QJ_ERR_INVALID;
  }
  MapArgs a;
  qj_status st = make_map_args(grid, f, params, &a);
  if (st) return st;
  const int64_t XY = (int64_t)a.res[0] * a.res[1];
  if (XY > kStaticSlabVoxels) {
    set_error("qj_field_eval_noalloc: a z-slice of %lld cells exceeds the maximum of %lld", (long long)XY,
              (long long)kStaticSlabVoxels);
    return QJ_ERR_CAPACITY;
  }
  int dev = 0;
  QJ_HIP(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  if ((st = scratch_setup(dev, &base))) return st;
  const FieldBufs fb{(uint32_t*)base, (uint32_t*)(base + kStaticListBytes), kStaticSlabVoxels, kStaticSlabVoxels,
                     (uint32_t*)(base + kStaticCountOff), (unsigned long long*)(base + kStaticCounterOff), nullptr};
  const hipStream_t s = (hipStream_t)stream;
  const int64_t Z = a.res[2], nz = kStaticSlabVoxels / XY;
  for (int64_t z0 = 0; z0 < Z; z0 += nz) {
    MapArgs sa = a;
    sa.res[2] = (int32_t)(Z - z0 < nz ? Z - z0 : nz);
    sa.zoff = (int32_t)z0;
    if ((st = launch_field_bufs(fb, sa, f, precision, band, d_field + z0 * XY, s))) return st;
  }
  return QJ_OK;
}
