// FIELD_3D geometry on the device.
#pragma once
#include <hip/hip_runtime.h>

namespace qj {

// FIELD_3D::cellCenter (FIELD_3D.cpp:818-841) in the reference's order with
// no contraction, so cell centres are bit-identical to the CPU path's. A is
// any struct with lo[3], dx[3], center[3], rot[9] (row-major rotation).
template <class A>
__device__ __forceinline__ void cell_center(const A& a, int x, int y, int z, double* out) {
#pragma clang fp contract(off)
  double f0 = a.lo[0], f1 = a.lo[1], f2 = a.lo[2];
  f0 += x * a.dx[0];
  f1 += y * a.dx[1];
  f2 += z * a.dx[2];
  f0 += a.dx[0] * 0.5;
  f1 += a.dx[1] * 0.5;
  f2 += a.dx[2] * 0.5;
  f0 -= a.center[0];
  f1 -= a.center[1];
  f2 -= a.center[2];
  const double r0 = a.rot[0] * f0 + a.rot[1] * f1 + a.rot[2] * f2;
  const double r1 = a.rot[3] * f0 + a.rot[4] * f1 + a.rot[5] * f2;
  const double r2 = a.rot[6] * f0 + a.rot[7] * f1 + a.rot[8] * f2;
  out[0] = r0 + a.center[0];
  out[1] = r1 + a.center[1];
  out[2] = r2 + a.center[2];
}

}  // namespace qj
