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

#include <vector>

#include "qj_geom.h"
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

// mean_at: [mean, count of summed cells]; when every cell was summed there is
// nothing to replace and the field is not read again
template <bool kNan>
__global__ __launch_bounds__(kOpsBlock) void k_replace(double* __restrict__ f, int64_t n,
                                                       const double* __restrict__ mean_at) {
  if (mean_at[1] == (double)n) return;  // the same for every thread of the grid
  const double mean = mean_at[0];
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

// ---------------------------------------------------------------- scoring fields of the shape optimizer

// outerShell + innerShell (FIELD_3D.cpp:5070-5145): 1 where the cell has a
// 26-neighbour of the other sign (cells at 0 have none)
__global__ __launch_bounds__(kOpsBlock) void k_shells(const double* __restrict__ d, int64_t X, int64_t Y, int64_t Z,
                                                      double* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x, XY = X * Y;
  if (i >= XY * Z) return;
  const int64_t z = i / XY, r = i - z * XY, y = r / X, x = r - y * X;
  const double c = d[i];
  bool flip = false;
  for (int k = -1; k <= 1; k++)
    for (int j = -1; j <= 1; j++)
      for (int h = -1; h <= 1; h++) {
        if (h == 0 && j == 0 && k == 0) continue;
        const int64_t xn = x + h, yn = y + j, zn = z + k;
        if (xn < 0 || xn > X - 1 || yn < 0 || yn > Y - 1 || zn < 0 || zn > Z - 1) continue;
        if (d[xn + yn * X + zn * XY] * c < 0) flip = true;
      }
  out[i] = flip ? 1.0 : 0.0;
}

// One sweep of invertedDistance's layer numbering (FIELD_3D.cpp:5248-5283):
// every cell at or below 1e-5 in `in` takes 1 + its smallest 6-neighbour
// above 1e-5, written to `out` (the sweep reads only the previous one, so the
// result does not depend on the order of the cells). cnt[0] counts the cells
// still at or below 1e-5 before the sweep, cnt[1] those the sweep numbered.
__global__ __launch_bounds__(kOpsBlock) void k_relax(const double* __restrict__ in, int64_t X, int64_t Y, int64_t Z,
                                                     double* __restrict__ out, unsigned long long* __restrict__ cnt) {
  __shared__ unsigned s_zero[kOpsBlock / 64], s_upd[kOpsBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x, XY = X * Y, n = XY * Z;
  bool zero = false, upd = false;
  if (i < n) {
    const double v = in[i];
    double w = v;
    if (!(v > 1e-5)) {
      zero = true;
      const int64_t z = i / XY, r = i - z * XY, y = r / X, x = r - y * X;
      double nb[6];
      nb[0] = (x != 0) ? in[i - 1] : 0;
      nb[1] = (x != X - 1) ? in[i + 1] : 0;
      nb[2] = (y != 0) ? in[i - X] : 0;
      nb[3] = (y != Y - 1) ? in[i + X] : 0;
      nb[4] = (z != 0) ? in[i - XY] : 0;
      nb[5] = (z != Z - 1) ? in[i + XY] : 0;
      double biggest = (double)n + 1;
      for (int k = 0; k < 6; k++)
        if (nb[k] > 1e-5 && nb[k] < biggest) biggest = nb[k];
      if (biggest < (double)n) {
        w = biggest + 1;
        upd = true;
      }
    }
    out[i] = w;
  }
  const unsigned long long mz = __ballot(zero), mu = __ballot(upd);
  if ((threadIdx.x & 63) == 0) {
    s_zero[threadIdx.x >> 6] = (unsigned)__popcll(mz);
    s_upd[threadIdx.x >> 6] = (unsigned)__popcll(mu);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (int k = 0; k < kOpsBlock / 64; k++) {
      a += s_zero[k];
      b += s_upd[k];
    }
    if (a) atomicAdd(&cnt[0], a);
    if (b) atomicAdd(&cnt[1], b);
  }
}

// 1 / value, negated inside (distance < 0) and quartered outside (FIELD_3D.cpp:5286-5296)
__global__ __launch_bounds__(kOpsBlock) void k_invert_signed(double* __restrict__ v, const double* __restrict__ d,
                                                             int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x;
  if (i >= n) return;
  double x = 1.0 / v[i];
  if (d[i] < 0) x *= -1;
  else x *= 0.25;
  v[i] = x;
}

// partial sums (double-double) of |v| (mode 0), of v where d < 0 (mode 1),
// of v where d >= 0 (mode 2)
__global__ __launch_bounds__(kOpsBlock) void k_masked_sum(const double* __restrict__ v, const double* __restrict__ d,
                                                          int64_t n, int mode, double* __restrict__ part) {
  Acc a{{0.0, 0.0}, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kOpsBlock;
  for (int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x; i < n; i += stride) {
    const double x = v[i];
    if (mode == 1 && !(d[i] < 0)) continue;
    if (mode == 2 && (d[i] < 0)) continue;
    const double y = mode == 0 ? fabs(x) : x;
    a.count += 1.0;
    if (isfinite(y)) a.s = dd_add1(a.s, y);
    else a.special += y;
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

// v *= inside-scale where d < 0, outside-scale elsewhere
__global__ __launch_bounds__(kOpsBlock) void k_scale_by_sign(double* __restrict__ v, const double* __restrict__ d,
                                                             int64_t n, double s_in, double s_out) {
  const int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x;
  if (i >= n) return;
  v[i] *= (d[i] < 0) ? s_in : s_out;
}

// weightScore / inverseDistanceScore (OPTIMIZE_3D.cpp:4502-4563) for threshold
// blockIdx.y: -tanh(alpha (F - threshold)) w(cell centre); the block for the
// last threshold slot also sums |w| (computeMaximumScaledPowerRationalScore,
// :4484-4498). w is the trilinear lookup of the weight field.
__global__ __launch_bounds__(kOpsBlock) void k_weight_score(const double* __restrict__ f, GridGeom g, GridGeom wg,
                                                            const double* __restrict__ w,
                                                            const double* __restrict__ thresholds, int nthr,
                                                            double alpha, double* __restrict__ part) {
  const int64_t n = (int64_t)g.res[0] * g.res[1] * g.res[2], X = g.res[0], XY = X * g.res[1];
  const int t = blockIdx.y;
  Acc a{{0.0, 0.0}, 0.0, 0.0};
  const int64_t stride = (int64_t)gridDim.x * kOpsBlock;
  for (int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x; i < n; i += stride) {
    const int64_t z = i / XY, r = i - z * XY, y = r / X, x = r - y * X;
    double c[3];
    cell_center(g, (int)x, (int)y, (int)z, c);
    const double wt = field_lookup(wg, w, c);
    const double v = t < nthr ? -tanh(alpha * (f[i] - thresholds[t])) * wt : fabs(wt);
    a.count += 1.0;
    if (isfinite(v)) a.s = dd_add1(a.s, v);
    else a.special += v;
  }
  a = block_reduce(a);
  if (threadIdx.x == 0) {
    double* p = part + 4 * ((int64_t)t * gridDim.x + blockIdx.x);
    p[0] = a.s.hi;
    p[1] = a.s.lo;
    p[2] = a.special;
    p[3] = a.count;
  }
}

// fieldMin / fieldMax (FIELD_3D.cpp:2960-2971, 3021-3032): the serial scan
// keeps the first cell that compares smaller (larger), starting from cell 0,
// so NaN cells are skipped unless cell 0 is NaN, and of equal values (0 and
// -0) the first wins. Reduced as (value, index) pairs with the index breaking
// ties; NaN cells never win.
struct VI {
  double v;
  int64_t i;
};
__device__ __forceinline__ VI vi_min(VI a, VI b) {
  if (b.i < 0) return a;
  if (a.i < 0) return b;
  return (b.v < a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}
__device__ __forceinline__ VI vi_max(VI a, VI b) {
  if (b.i < 0) return a;
  if (a.i < 0) return b;
  return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}

__global__ __launch_bounds__(kOpsBlock) void k_min_max(const double* __restrict__ f, int64_t n,
                                                       double* __restrict__ part) {
  __shared__ double s_v[2][kOpsBlock];
  __shared__ int64_t s_i[2][kOpsBlock];
  VI mn{0.0, -1}, mx{0.0, -1};
  const int64_t stride = (int64_t)gridDim.x * kOpsBlock;
  for (int64_t i = (int64_t)blockIdx.x * kOpsBlock + threadIdx.x; i < n; i += stride) {
    const double v = f[i];
    if (isnan(v)) continue;
    mn = vi_min(mn, VI{v, i});
    mx = vi_max(mx, VI{v, i});
  }
  const int t = threadIdx.x;
  s_v[0][t] = mn.v;
  s_i[0][t] = mn.i;
  s_v[1][t] = mx.v;
  s_i[1][t] = mx.i;
  __syncthreads();
  for (int w = kOpsBlock / 2; w > 0; w >>= 1) {
    if (t < w) {
      const VI a = vi_min(VI{s_v[0][t], s_i[0][t]}, VI{s_v[0][t + w], s_i[0][t + w]});
      const VI b = vi_max(VI{s_v[1][t], s_i[1][t]}, VI{s_v[1][t + w], s_i[1][t + w]});
      s_v[0][t] = a.v;
      s_i[0][t] = a.i;
      s_v[1][t] = b.v;
      s_i[1][t] = b.i;
    }
    __syncthreads();
  }
  if (t == 0) {
    double* p = part + 4 * (int64_t)blockIdx.x;
    p[0] = s_v[0][0];
    p[1] = (double)s_i[0][0];  // exact: indexes stay below 2^53
    p[2] = s_v[1][0];
    p[3] = (double)s_i[1][0];
  }
}

// host: the sum of nparts partials (4 doubles each) read back
static double host_sum(const std::vector<double>& p, int64_t first, int64_t nparts) {
  double hi = 0, lo = 0, special = 0;
  for (int64_t k = first; k < first + nparts; k++) {
    // two-sum on the host too, so the total stays the correctly rounded sum
    const double a = p[4 * k], b = p[4 * k + 1];
    double s = hi + a, bb = s - hi, e = (hi - (s - bb)) + (a - bb);
    lo += e + b;
    hi = s;
    special += p[4 * k + 2];
  }
  return isfinite(special) ? hi + lo : special;
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

namespace qj {

qj_status field_inverted_distance(qj_workspace* ws, const double* d_dist, const qj_grid* grid, bool normalized,
                                  double* d_out, double* abs_sum, hipStream_t s) {
  const int64_t X = grid->res[0], Y = grid->res[1], Z = grid->res[2], n = X * Y * Z;
  if (n <= 0) return QJ_OK;
  const unsigned nb = (unsigned)((n + kOpsBlock - 1) / kOpsBlock);
  double* tmp = nullptr;
  unsigned long long* cnt = nullptr;
  QJ_HIP(hipMalloc(&tmp, sizeof(double) * (size_t)n));
  struct Free {
    void* a;
    void* b;
    ~Free() {
      (void)hipFree(a);
      (void)hipFree(b);
    }
  } guard{tmp, nullptr};
  QJ_HIP(hipMalloc(&cnt, 2 * sizeof(unsigned long long)));
  guard.b = cnt;
  hipLaunchKernelGGL(k_shells, dim3(nb), dim3(kOpsBlock), 0, s, d_dist, X, Y, Z, d_out);
  QJ_HIP(hipGetLastError());
  // sweeps until no cell is left at or below 1e-5; a sweep that numbers no
  // cell while some remain would repeat forever in the reference
  double* cur = d_out;
  double* nxt = tmp;
  for (;;) {
    unsigned long long h[2] = {0, 0};
    QJ_HIP(hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_relax, dim3(nb), dim3(kOpsBlock), 0, s, (const double*)cur, X, Y, Z, nxt, cnt);
    QJ_HIP(hipGetLastError());
    QJ_HIP(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
    std::swap(cur, nxt);
    if (h[0] == 0) break;
    if (h[1] == 0) {
      set_error("inverted distance: %llu cells can never be numbered (no surface reaches them)", h[0]);
      return QJ_ERR_NUMERIC;
    }
  }
  if (cur != d_out) QJ_HIP(hipMemcpyAsync(d_out, cur, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_invert_signed, dim3(nb), dim3(kOpsBlock), 0, s, d_out, d_dist, n);
  QJ_HIP(hipGetLastError());
  if (!normalized && !abs_sum) return QJ_OK;
  qj_status st = ensure_partials(ws, 4 * (int64_t)kOpsGrid, s);
  if (st) return st;
  const int grid_sum = (int)(nb < (unsigned)kOpsGrid ? nb : kOpsGrid);
  std::vector<double> part(4 * (size_t)grid_sum);
  auto sum = [&](int mode, double* out) -> qj_status {
    hipLaunchKernelGGL(k_masked_sum, dim3(grid_sum), dim3(kOpsBlock), 0, s, (const double*)d_out, d_dist, n, mode,
                       ws->d_partials);
    QJ_HIP(hipGetLastError());
    QJ_HIP(hipMemcpyAsync(part.data(), ws->d_partials, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
    QJ_HIP(hipStreamSynchronize(s));
    *out = host_sum(part, 0, grid_sum);
    return QJ_OK;
  };
  if (normalized) {
    // invertedDistanceNormalized (FIELD_3D.cpp:5316-5352)
    double total = 0, inside = 0, outside = 0;
    if ((st = sum(0, &total))) return st;
    const double scale = 1.0 / total;
    hipLaunchKernelGGL(k_scale_by_sign, dim3(nb), dim3(kOpsBlock), 0, s, d_out, d_dist, n, scale, scale);
    QJ_HIP(hipGetLastError());
    if ((st = sum(1, &inside)) || (st = sum(2, &outside))) return st;
    hipLaunchKernelGGL(k_scale_by_sign, dim3(nb), dim3(kOpsBlock), 0, s, d_out, d_dist, n, 0.5 / fabs(inside),
                       0.5 / fabs(outside));
    QJ_HIP(hipGetLastError());
  }
  if (abs_sum && (st = sum(0, abs_sum))) return st;
  return QJ_OK;
}

qj_status field_min_max(qj_workspace* ws, const double* d_field, int64_t n, double* mn, double* mx, hipStream_t s) {
  if (n <= 0) return QJ_ERR_INVALID;
  const int64_t want = (n + kOpsBlock - 1) / kOpsBlock;
  const int grid = (int)(want < kOpsGrid ? want : kOpsGrid);
  qj_status st = ensure_partials(ws, 4 * (int64_t)grid, s);
  if (st) return st;
  hipLaunchKernelGGL(k_min_max, dim3(grid), dim3(kOpsBlock), 0, s, d_field, n, ws->d_partials);
  QJ_HIP(hipGetLastError());
  std::vector<double> part(4 * (size_t)grid);
  double first = 0;
  QJ_HIP(hipMemcpyAsync(part.data(), ws->d_partials, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipMemcpyAsync(&first, d_field, sizeof(double), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipStreamSynchronize(s));
  if (isnan(first)) {  // nothing compares against a NaN start
    *mn = *mx = first;
    return QJ_OK;
  }
  double bv = 0, tv = 0;
  int64_t bi = -1, ti = -1;
  for (int b = 0; b < grid; b++) {
    const int64_t i0 = (int64_t)part[4 * b + 1], i1 = (int64_t)part[4 * b + 3];
    const double v0 = part[4 * b], v1 = part[4 * b + 2];
    if (i0 >= 0 && (bi < 0 || v0 < bv || (v0 == bv && i0 < bi))) {
      bv = v0;
      bi = i0;
    }
    if (i1 >= 0 && (ti < 0 || v1 > tv || (v1 == tv && i1 < ti))) {
      tv = v1;
      ti = i1;
    }
  }
  *mn = bv;
  *mx = tv;
  return QJ_OK;
}

qj_status weight_scores(qj_workspace* ws, const double* d_field, const qj_grid* grid, const qj_grid* wgrid,
                        const double* d_weight, double alpha, const double* thresholds, int nthr, double* scores,
                        double* max_score, hipStream_t s) {
  GridGeom g, wg;
  make_geom(grid, &g);
  make_geom(wgrid, &wg);
  const int64_t n = (int64_t)g.res[0] * g.res[1] * g.res[2];
  if (n <= 0 || nthr < 0) return QJ_ERR_INVALID;
  const int64_t want = (n + kOpsBlock - 1) / kOpsBlock;
  const int gx = (int)(want < 256 ? want : 256);
  const int slots = nthr + 1;
  qj_status st = ensure_partials(ws, 4 * (int64_t)gx * slots + (int64_t)slots, s);
  if (st) return st;
  double* d_thr = ws->d_partials + 4 * (int64_t)gx * slots;
  if (nthr) QJ_HIP(hipMemcpyAsync(d_thr, thresholds, sizeof(double) * (size_t)nthr, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_weight_score, dim3(gx, slots), dim3(kOpsBlock), 0, s, d_field, g, wg, d_weight,
                     (const double*)d_thr, nthr, alpha, ws->d_partials);
  QJ_HIP(hipGetLastError());
  std::vector<double> part(4 * (size_t)gx * slots);
  QJ_HIP(hipMemcpyAsync(part.data(), ws->d_partials, sizeof(double) * part.size(), hipMemcpyDeviceToHost, s));
  QJ_HIP(hipStreamSynchronize(s));
  for (int t = 0; t < nthr; t++) scores[t] = host_sum(part, (int64_t)t * gx, gx);
  if (max_score) *max_score = host_sum(part, (int64_t)nthr * gx, gx);
  return QJ_OK;
}

}  // namespace qj
