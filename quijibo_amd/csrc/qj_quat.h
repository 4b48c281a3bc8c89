// Device quaternion algebra for the Julia field kernels (gfx950).
//
// Slot order is the reference's QUATERNION(w, x, y, z) (QUATERNION.h:36); w is
// the real part. The FP64 routines evaluate in the reference's expression
// order (QUATERNION.cpp:154-200, 375-393) so the FP64 kernels track the CPU
// path to a few ulp per operation; only fused multiply-add contraction and the
// device math library differ.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qj {

struct q64 {
  double w, x, y, z;
};
struct q32 {
  float w, x, y, z;
};

// operator*(QUATERNION, QUATERNION): QUATERNION.cpp:154-162
__device__ __forceinline__ q64 qmul(const q64& l, const q64& r) {
  q64 f;
  f.x = l.y * r.z - l.z * r.y + r.w * l.x + l.w * r.x;
  f.y = l.z * r.x - l.x * r.z + r.w * l.y + l.w * r.y;
  f.z = l.x * r.y - l.y * r.x + r.w * l.z + l.w * r.z;
  f.w = l.w * r.w - l.x * r.x - r.y * l.y - l.z * r.z;
  return f;
}
__device__ __forceinline__ q32 qmul(const q32& l, const q32& r) {
  q32 f;
  f.x = l.y * r.z - l.z * r.y + r.w * l.x + l.w * r.x;
  f.y = l.z * r.x - l.x * r.z + r.w * l.y + l.w * r.y;
  f.z = l.x * r.y - l.y * r.x + r.w * l.z + l.w * r.z;
  f.w = l.w * r.w - l.x * r.x - r.y * l.y - l.z * r.z;
  return f;
}

__device__ __forceinline__ q64 qscale(const q64& q, double s) { return {q.w * s, q.x * s, q.y * s, q.z * s}; }
__device__ __forceinline__ q32 qscale(const q32& q, float s) { return {q.w * s, q.x * s, q.y * s, q.z * s}; }
__device__ __forceinline__ q64 qsub(const q64& a, const q64& b) { return {a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ q32 qsub(const q32& a, const q32& b) { return {a.w - b.w, a.x - b.x, a.y - b.y, a.z - b.z}; }

// inverse(): conjugate() * (1.0 / |r|^2), QUATERNION.cpp:167-172, 196-200
__device__ __forceinline__ q64 qinv(const q64& r) {
  const double m = r.w * r.w + r.x * r.x + r.y * r.y + r.z * r.z;
  const double inv = 1.0 / m;
  return {r.w * inv, -r.x * inv, -r.y * inv, -r.z * inv};
}
// operator/(QUATERNION, QUATERNION): left * inverse(right), QUATERNION.cpp:174-182
__device__ __forceinline__ q64 qdiv(const q64& l, const q64& r) { return qmul(l, qinv(r)); }

__device__ __forceinline__ double qmag(const q64& q) { return sqrt(q.w * q.w + q.x * q.x + q.y * q.y + q.z * q.z); }
__device__ __forceinline__ double qdot(const q64& a, const q64& b) { return a.w * b.w + a.x * b.x + a.y * b.y + a.z * b.z; }

// QUATERNION::pow, QUATERNION.cpp:375-393, including its guards: a purely
// real base gives (+|q|^e, 0, 0, 0); a non-positive angle e*theta drops the
// imaginary part; q = 0 gives NaN.
__device__ __forceinline__ q64 qpow(const q64& q, double e) {
  const double partial = q.x * q.x + q.y * q.y + q.z * q.z;
  const double qMagnitude = sqrt(partial + q.w * q.w);
  const double vMagnitude = sqrt(partial);
  const double vMagnitudeInv = (vMagnitude > 0.0) ? 1.0 / vMagnitude : 0.0;
  const double scale = e * acos(q.w / qMagnitude) * vMagnitudeInv;
  const double magnitude = scale * vMagnitude;
  const double magnitudeInv = (magnitude > 0.0) ? 1.0 / magnitude : 0.0;
  const double exps = exp(e * log(qMagnitude));
  double s, c;
  sincos(magnitude, &s, &c);
  const double scale2 = scale * exps * magnitudeInv * s;
  return {exps * c, scale2 * q.x, scale2 * q.y, scale2 * q.z};
}

// QUATERNION::log, QUATERNION.cpp:352-367. Its vector magnitude goes through
// TVEC3::magnitude, which is sqrtf of the double sum (VEC3.h:101-103).
__device__ __forceinline__ q64 qlog(const q64& q) {
  const double qMagnitude = qmag(q);
  const double vMagnitude = (double)__fsqrt_rn((float)(q.x * q.x + q.y * q.y + q.z * q.z));
  double nx = 0, ny = 0, nz = 0;
  if (vMagnitude > 0) {
    nx = q.x / vMagnitude;
    ny = q.y / vMagnitude;
    nz = q.z / vMagnitude;
  }
  const double a = acos(q.w / qMagnitude);
  return {log(qMagnitude), nx * a, ny * a, nz * a};
}

}  // namespace qj
