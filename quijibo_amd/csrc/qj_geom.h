// FIELD_3D geometry on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

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

// FIELD_3D::operator()(VEC3F) (FIELD_3D.cpp:966-1029): trilinear with float weights
// G: any struct with res[3], dx[3], center[3] and lengths[3] (GridGeom).
template <class G>
__device__ __forceinline__ double field_lookup(const G& g, const double* __restrict__ d, const double* p) {
#pragma clang fp contract(off)
  const int xRes = g.res[0], yRes = g.res[1], zRes = g.res[2];
  const double invDx = 1.0 / g.dx[0], invDy = 1.0 / g.dx[1], invDz = 1.0 / g.dx[2];
  double c0 = g.center[0] - 0.5 * g.lengths[0], c1 = g.center[1] - 0.5 * g.lengths[1],
         c2 = g.center[2] - 0.5 * g.lengths[2];
  c0 += 0.5 * g.dx[0];
  c1 += 0.5 * g.dx[1];
  c2 += 0.5 * g.dx[2];
  double p0 = p[0] - c0, p1 = p[1] - c1, p2 = p[2] - c2;
  p0 *= invDx;
  p1 *= invDy;
  p2 *= invDz;
  int x0 = (int)p0, y0 = (int)p1, z0 = (int)p2;
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  x0 = x0 < 0 ? 0 : x0; y0 = y0 < 0 ? 0 : y0; z0 = z0 < 0 ? 0 : z0;
  x1 = x1 < 0 ? 0 : x1; y1 = y1 < 0 ? 0 : y1; z1 = z1 < 0 ? 0 : z1;
  x0 = x0 > xRes - 1 ? xRes - 1 : x0; y0 = y0 > yRes - 1 ? yRes - 1 : y0; z0 = z0 > zRes - 1 ? zRes - 1 : z0;
  x1 = x1 > xRes - 1 ? xRes - 1 : x1; y1 = y1 > yRes - 1 ? yRes - 1 : y1; z1 = z1 > zRes - 1 ? zRes - 1 : z1;
  const float s1 = (float)(p0 - x0), s0 = 1.0f - s1;
  const float t1 = (float)(p1 - y0), t0 = 1.0f - t1;
  const float u1 = (float)(p2 - z0), u0 = 1.0f - u1;
  const int64_t slab = (int64_t)xRes * yRes;
  const double d000 = d[x0 + y0 * xRes + z0 * slab], d010 = d[x0 + y1 * xRes + z0 * slab];
  const double d100 = d[x1 + y0 * xRes + z0 * slab], d110 = d[x1 + y1 * xRes + z0 * slab];
  const double d001 = d[x0 + y0 * xRes + z1 * slab], d011 = d[x0 + y1 * xRes + z1 * slab];
  const double d101 = d[x1 + y0 * xRes + z1 * slab], d111 = d[x1 + y1 * xRes + z1 * slab];
  return u0 * (s0 * (t0 * d000 + t1 * d010) + s1 * (t0 * d100 + t1 * d110)) +
         u1 * (s0 * (t0 * d001 + t1 * d011) + s1 * (t0 * d101 + t1 * d111));
}

}  // namespace qj
