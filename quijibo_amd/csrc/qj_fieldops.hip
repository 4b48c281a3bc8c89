// FIELD_3D clean-up passes that the reference runs on a fractal before
// meshing it (toothOptimizer.cpp:250-254, OPTIMIZE_3D.cpp:4718):
//   clampInfsToMean  (FIELD_3D.cpp:5182-5207)  every infinite cell takes the
//                    mean of the non-infinite cells
//   clampNansToMean  (FIELD_3D.cpp:5212-5234)  every NaN cell takes the mean
//                    of the non-NaN cells
//   setBorders       (FIELD_3D.cpp:5357-5382)  the six faces take a constant
//
// The reference sums serially in x-fastest order, so a NaN cell makes the
// clampInfsToMean mean NaN and an infinite cell makes the clampNansToMean
// mean infinite (or NaN with both signs). Here the finite cells are summed in
// double-double arithmetic (error-free TwoSum per addition, a fixed tree over
// blocks, so the result does not depend on scheduling) and the non-finite
// ones separately in plain IEEE arithmetic, which gives the serial sum's
// non-finite outcomes exactly. The double-double sum is off the exact sum by
// at most about n u^2 sum|x_i| before its final rounding (u = 2^-53), so the
// mean is the correctly rounded sum over the count for any field that fits
// in memory; the reference's serial sum differs from that by up to
// (n - 1) u sum|x_i|. A finite sum beyond DBL_MAX gives NaN here (the
// serial sum gives inf).
//
// HBM-bound: the field is read twice (sum, then replace) and only replaced
// cells are written; the mean never leaves the device, so nothing waits on
// the host. Three launches: per-block partial sums, one block that combines
// them into the mean, and the replacement.
#include <math.h>

#include "qj_internal.h"

namespace qj {

namespace {

constexpr int kOpsBlock = 256;
constexpr int kOpsGrid = 2048;  // 8 blocks per CU; a fixed count keeps the reduction tree fixed

struct DD {
  double hi, lo;
};

// Knuth's TwoSum: s + e == a + b exactly (no multiplications, so no contraction applies).
__device__ __forceinline__ DD two_sum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  const double e = (a - (s - bb)) + (b - bb);
  return DD{s, e};
}
__device__ __forceinline__ DD fast_two_sum(double a, double b) {
  const double s = a + b;
  return DD{s, b - (s - a)};
}
__device__ __forceinline__ DD dd_add(DD a, DD b) {
  DD s = two_sum(a.hi, b.hi);
  s.lo += a.lo + b.lo;
  return fast_two_sum(s.hi, s.lo);
}
__device__ __forceinline__ DD dd_add1(DD a, double x) {
  DD s = two_sum(a.hi, x);
  s.lo += a.lo;
  return fast_two_sum(s.hi, s.lo);
}

// kNan false: cells summed are the non-infinite ones (clampInfsToMean);
// true: the non-NaN ones (clampNansToMean).
template <bool kNan>
__device__ __forceinline__ bool summed(double v) {
  return kNan ? !isnan(v) : !isinf(v);
}

struct Acc {
  DD s;
  double special;  // sum of the summed non-finite cells: 0, +-inf or NaN
  double count;    // exact below 2^53 cells
};

__device__ __forceinline__ Acc acc_add(Acc a, const Acc& b) {
  a.s = dd_add(a.s, b.s);
  a.special += b.special;
  a.count += b.count;
  return a;
}

// block reduction in a fixed order (tree over the 256 lanes through LDS)
__device__ Acc block_reduce(Acc a) {
  __shared__ double s_hi[kOpsBlock], s_lo[kOpsBlock], s_sp[kOpsBlock], s_ct[kOpsBlock];
  const int t = threadIdx.x;
  s_hi[t] = a.s.hi;
  s_lo[t] = a.s.lo;
  s_sp[t] = a.special;
  s_ct[t] = a.count;
  __syncthreads();
  for (int w = kOpsBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      Acc x{{s_hi[t], s_lo[t]}, s_sp[t], s_ct[t]};
      const Acc y{{s_hi[t + w], s_lo[t + w]}, s_sp[t + w], s_ct[t + w]};
      x = acc_add(x, y);
      s_hi[t] = x.s.hi;
      s_lo[t] = x.s.lo;
      s_sp[t] = x.special;
      s_ct[t] = x.count;
    }
    __syncthreads();
  }
  return Acc{{s_hi[0], s_lo[0]}, s_sp[0], s_ct[0]};
}

template <bool kNan>
__global__ __launch_bounds__(kOpsBlock) void k_mean_partials(const double* __restrict__ f, int64_t n,
                                                            double* __restrict__ part) {
  Acc a{{0.0, 0.0}, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kOpsBlock;
  for (int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x; i < n; i += stride) {
    const double v = f[i];
    if (summed<kNan>(v)) {
      a.count += 1.0;
      if (isfinite(v)) a.s = dd_add1(a.s, v);
      else a.special += v;
    }
  }
  a = block_reduce(a);
  if (threadIdx.x == 0) {
    double* p = part + 4 * (int64_t)blockIdx.x;
    p[0] = a.s.hi;
    p[1] = a.s.lo;
    p[2] = a.special;
    p[3] = a.count;
  }
}

// one block: combine the partials, write [mean, count] after them
__global__ __launch_bounds__(kOpsBlock) void k_mean_final(double* __restrict__ part, int nparts) {
  Acc a{{0.0, 0.0}, 0.0, 0.0};
  for (int i = threadIdx.x; i < nparts; i += kOpsBlock) {
    const double* p = part + 4 * (int64_t)i;
    a = acc_add(a, Acc{{p[0], p[1]}, p[2], p[3]});
  }
  a = block_reduce(a);
  if (threadIdx.x == 0) {
    // Real mean = sum / nonInfsSeen: 0 / 0 is NaN when no cell is summed
    const double total = isfinite(a.special) ? a.s.hi + a.s.lo : a.special;
    part[4 * (int64_t)nparts] = total / a.count;
    part[4 * (int64_t)nparts + 1] = a.count;
  }
}

template <bool kNan>
__global__ __launch_bounds__(kOpsBlock) void k_replace(double* __restrict__ f, int64_t n,
                                                       const double* __restrict__ mean_at) {
  const double mean = *mean_at;
  const int64_t stride = (int64_t)gridDim.x * kOpsBlock;
  for (int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x; i < n; i += stride) {
    const double v = f[i];
    if (kNan ? isnan(v) : isinf(v)) f[i] = mean;
  }
}

// the six faces: z = 0, Z-1 (X*Y cells each), y = 0, Y-1 (X*Z), x = 0, X-1 (Y*Z)
__global__ __launch_bounds__(kOpsBlock) void k_set_borders(double* __restrict__ f, int64_t X, int64_t Y, int64_t Z,
                                                           double value) {
  const int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x;
  const int64_t nz = X * Y, ny = X * Z, nx = Y * Z;
  int64_t x, y, z;
  if (i < 2 * nz) {
    const int64_t r = i % nz;
    x = r % X;
    y = r / X;
    z = i < nz ? 0 : Z - 1;
  } else if (i < 2 * (nz + ny)) {
    const int64_t j = i - 2 * nz, r = j % ny;
    x = r % X;
    z = r / X;
    y = j < ny ? 0 : Y - 1;
  } else if (i < 2 * (nz + ny + nx)) {
    const int64_t j = i - 2 * (nz + ny), r = j % nx;
    y = r % Y;
    z = r / Y;
    x = j < nx ? 0 : X - 1;
  } else {
    return;
  }
  f[x + y * X + z * X * Y] = value;
}

}  // namespace

qj_status field_clamp_to_mean(qj_workspace* ws, double* d_field, int64_t n, bool nans, double* mean_out,
                              hipStream_t s) {
  if (n <= 0) {
    if (mean_out) *mean_out = NAN;  // 0 / 0: no cell is summed
    return QJ_OK;
  }
  qj_status st = ensure_partials(ws, 4 * (int64_t)kOpsGrid + 2, s);
  if (st) return st;
  const int64_t want = (n + kOpsBlock - 1) / kOpsBlock;
  const int grid = (int)(want < kOpsGrid ? want : kOpsGrid);
  if (nans) hipLaunchKernelGGL(k_mean_partials<true>, dim3(grid), dim3(kOpsBlock), 0, s, d_field, n, ws->d_partials);
  else hipLaunchKernelGGL(k_mean_partials<false>, dim3(grid), dim3(kOpsBlock), 0, s, d_field, n, ws->d_partials);
  QJ_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(kOpsBlock), 0, s, ws->d_partials, grid);
  QJ_HIP(hipGetLastError());
  const double* mean_at = ws->d_partials + 4 * (int64_t)grid;
  if (nans) hipLaunchKernelGGL(k_replace<true>, dim3(grid), dim3(kOpsBlock), 0, s, d_field, n, mean_at);
  else hipLaunchKernelGGL(k_replace<false>, dim3(grid), dim3(kOpsBlock), 0, s, d_field, n, mean_at);
  QJ_HIP(hipGetLastError());
  if (mean_out) {
    QJ_HIP(hipMemcpyAsync(mean_out, mean_at, sizeof(double), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
  }
  return QJ_OK;
}

qj_status field_set_borders(const int32_t* res, double value, double* d_field, hipStream_t s) {
  const int64_t X = res[0], Y = res[1], Z = res[2];
  const int64_t m = 2 * (X * Y + X * Z + Y * Z);
  const int64_t blocks = (m + kOpsBlock - 1) / kOpsBlock;
  if (blocks > INT32_MAX) return QJ_ERR_CAPACITY;
  hipLaunchKernelGGL(k_set_borders, dim3((unsigned)blocks), dim3(kOpsBlock), 0, s, d_field, X, Y, Z, value);
  QJ_HIP(hipGetLastError());
  return QJ_OK;
}

}  // namespace qj
