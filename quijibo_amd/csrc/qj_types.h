// Types shared by the kernels and the host layer.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qj_c_api.h"

namespace qj {

// Per-root records in HBM. Top roots first, then bottom roots.
struct RootRec64 {
  double w, x, y, z;
  double e;   // power_scalar * powers[i], rounded as the reference does (POLYNOMIAL_4D.cpp:914,919)
  double p;   // powers[i]
};
// FP32 view of one root (host side; kernels read the packed block below).
struct RootRec32 {
  float w, x, y, z;
  float e;
  int32_t kind;  // 1: e == 1, 2: e == 2 (exact integer powers), 0: general power
};

struct alignas(16) F4 {
  float w, x, y, z;
};

// Device block of the FP32 roots of one rational function, top roots then
// bottom roots (n = nt + nb): F4 q[n] | int32 pat[n] | float e[n]. pat[i]
// holds the exponent kinds of roots i..i+3 of the same polynomial, two bits
// per root (bit 2j: root i+j has e = 2; bit 2j+1: its exponent is neither 1
// nor 2), so a group of four roots is one 64-byte scalar load and one word.
constexpr int kPatSquare = 1, kPatGeneral = 2;
struct Roots32 {
  const F4* q;
  const int32_t* pat;
  const float* e;
};
__host__ __device__ inline Roots32 roots32_view(const void* blk, int n) {
  const F4* q = (const F4*)blk;
  const int32_t* pat = (const int32_t*)(q + n);
  return Roots32{q, pat, (const float*)(pat + n)};
}
// Geometry and map constants passed by value to every kernel.
struct MapArgs {
  int32_t res[3];
  int32_t max_iter;
  int32_t zoff;      // first z slice of the evaluated range within the grid (cell centres use z + zoff)
  int32_t gen;               // 1 when some root's exponent is not exactly 1 or 2 (FP32 general kind)
  double lo[3];      // center - 0.5 * lengths (FIELD_3D.cpp:822-825)
  double dx[3];      // lengths / res (FIELD_3D.cpp:50-52)
  double center[3];
  double rot[9];     // QUATERNION::toRotationMatrix of the grid rotation (QUATERNION.cpp:395-412)
  double slice;
  double exp_scaling;
  double escape;
  uint32_t flags;
  int32_t nt, nb;
  double sc_top, sc_bottom;  // power scalars
  double term_bits;          // max over roots of |e| log2(escape + |root|) + 2: bound on one term's log2 size
  double near_bits;          // max |e| log2(2 max|root| + 1) + 1: bound on a term's log2 size when q is within 1 of a root
};

// Per-map constants of the FP32 kernels (log2 units).
struct Consts32 {
  float log2c;      // log2(exp_scaling)
  float log2esc;    // log2(escape)
  float log2bh;     // log2(10 FLT_MIN)
  float margin;     // decision margin in log2 units
  float band;
  float term_bits;  // bound on log2 |term| between renormalisations
};
// Counter slots (summed over QJ_NCOUNTER_SPREAD copies to spread atomics).
enum {
  CNT_VOXELS = 0,
  CNT_ITERS = 1,
  CNT_BAND_VOXELS = 2,
  CNT_BAND_ITERS = 3,
  CNT_ADAPTIVE = 4,
  CNT_NONFINITE = 5,
  CNT_CHECKED = 6,        // cells re-evaluated by the sampled FP64 cross-check
  CNT_CHECK_FLIPS = 7,    // of those, sign disagreements of (float)F > 0
  CNT_CHECK_CLASS = 8,    // finite / NaN / inf class disagreements
  CNT_CHECK_MAXABS = 9,   // max |dF| (bits of a non-negative double, atomic max)
  CNT_CHECK_MAXREL = 10,  // max |dF| / max(1, |F64|)
  CNT_N = 12
};
constexpr int kCounterSpread = 64;

}  // namespace qj
