// C++ host layer: the reference's class API (POLYNOMIAL_4D, FIELD_3D,
// TRIANGLE_MESH, OPTIMIZE_3D) over the qj_* C ABI. Field work runs on the
// GPU; the classes keep the reference's host-memory ownership, so every call
// copies its inputs to HBM and its results back.
#include "../../include/quijibo/quijibo.hpp"

#include <hip/hip_runtime_api.h>
#include <float.h>
#include <math.h>
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <mutex>

#pragma STDC FP_CONTRACT OFF

namespace quijibo {

void check(qj_status s, const char* where) {
  if (s == QJ_OK) return;
  throw Error(s, std::string(where) + ": " + qj_status_string(s) + " (" + qj_last_error() + ")");
}

namespace {

void hip_check(hipError_t e, const char* where) {
  if (e != hipSuccess) throw Error(QJ_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

// One workspace per device for the classes that have none of their own.
qj_workspace* shared_workspace() {
  static std::mutex mu;
  static std::vector<qj_workspace*> per_device;
  int dev = 0;
  hip_check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lock(mu);
  if ((int)per_device.size() <= dev) per_device.resize(dev + 1, nullptr);
  if (!per_device[dev]) check(qj_workspace_create(&per_device[dev]), "qj_workspace_create");
  return per_device[dev];
}

// Device buffer owned for the duration of one call.
struct DeviceBuffer {
  void* p = nullptr;
  explicit DeviceBuffer(size_t bytes) {
    if (bytes) hip_check(hipMalloc(&p, bytes), "hipMalloc");
  }
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
};

// Byte sinks/sources for the FILE* and gz variants of the same layouts.
struct Writer {
  FILE* f = nullptr;
  gzFile g = nullptr;
  void put(const void* p, size_t n) {
    if (!n) return;
    if (g ? gzwrite(g, p, (unsigned)n) != (int)n : fwrite(p, 1, n, f) != n)
      throw Error(QJ_ERR_INVALID, "write failed");
  }
  void i32(int v) { put(&v, sizeof v); }
  void f64(double v) { put(&v, sizeof v); }
};

struct Reader {
  FILE* f = nullptr;
  gzFile g = nullptr;
  void get(void* p, size_t n) {
    if (!n) return;
    if (g ? gzread(g, p, (unsigned)n) != (int)n : fread(p, 1, n, f) != n)
      throw Error(QJ_ERR_INVALID, "read failed: file too short");
  }
  int i32() {
    int v;
    get(&v, sizeof v);
    return v;
  }
  double f64() {
    double v;
    get(&v, sizeof v);
    return v;
  }
};

void write_vec3(Writer& w, const VEC3F& v) {
  for (int i = 0; i < 3; i++) w.f64(v[i]);
}
VEC3F read_vec3(Reader& r) {
  VEC3F v;
  for (int i = 0; i < 3; i++) v[i] = r.f64();
  return v;
}

void write_field(Writer& w, const FIELD_3D& f) {
  w.i32(f.xRes());
  w.i32(f.yRes());
  w.i32(f.zRes());
  write_vec3(w, f.center());
  write_vec3(w, f.lengths());
  if (f.totalCells() > 0) w.put(f.dataConst(), sizeof(double) * f.totalCells());
}

void write_poly(Writer& w, const POLYNOMIAL_4D& p) {
  // POLYNOMIAL_4D::write (POLYNOMIAL_4D.cpp:650-685): totalRoots, then the
  // expanded coefficient, derivative and second-derivative arrays (written
  // empty: the factored map never uses them), roots, powers, powerScalar
  w.i32(p.totalRoots());
  w.i32(0);
  w.i32(0);
  w.i32(0);
  w.i32(p.totalRoots());
  for (const QUATERNION& q : p.roots())
    for (int i = 0; i < 4; i++) w.f64(q[i]);
  w.i32((int)p.powers().size());
  for (Real e : p.powers()) w.f64(e);
  w.f64(p.powerScalar());
}

POLYNOMIAL_4D read_poly(Reader& r) {
  // POLYNOMIAL_4D::read (POLYNOMIAL_4D.cpp:690-728): the coefficient arrays
  // are skipped; as in the reference, totalRoots powers are read
  const int total = r.i32();
  if (total < 0) throw Error(QJ_ERR_INVALID, "POLYNOMIAL_4D: negative root count");
  for (int k = 0; k < 3; k++) {
    const int n = r.i32();
    if (n < 0) throw Error(QJ_ERR_INVALID, "POLYNOMIAL_4D: negative array size");
    std::vector<double> skip(4 * (size_t)n);
    r.get(skip.data(), skip.size() * sizeof(double));
  }
  const int nr = r.i32();
  if (nr != total) throw Error(QJ_ERR_INVALID, "POLYNOMIAL_4D: root count mismatch");
  std::vector<QUATERNION> roots(nr);
  for (QUATERNION& q : roots)
    for (int i = 0; i < 4; i++) q[i] = r.f64();
  (void)r.i32();  // stored power count; totalRoots powers follow
  std::vector<Real> powers(total);
  for (Real& e : powers) e = r.f64();
  POLYNOMIAL_4D p(roots, powers);
  p.powerScalar() = r.f64();
  return p;
}

FIELD_3D read_field(Reader& r) {
  const int x = r.i32(), y = r.i32(), z = r.i32();
  const VEC3F c = read_vec3(r), l = read_vec3(r);
  if (x < 0 || y < 0 || z < 0) return FIELD_3D();
  FIELD_3D f(x, y, z, c, l);
  if (f.totalCells() > 0) r.get(f.data(), sizeof(double) * f.totalCells());
  return f;
}

bool ends_with_gz(const std::string& s) {
  return s.size() >= 2 && s[s.size() - 1] == 'z' && s[s.size() - 2] == 'g';
}

}  // namespace

// ---------------- VEC3F / QUATERNION ----------------

void VEC3F::write(FILE* f) const {
  Writer w{f};
  write_vec3(w, *this);
}
void VEC3F::read(FILE* f) {
  Reader r{f};
  *this = read_vec3(r);
}
void QUATERNION::write(FILE* f) const {
  Writer w{f};
  for (int i = 0; i < 4; i++) w.f64((*this)[i]);
}
void QUATERNION::read(FILE* f) {
  Reader r{f};
  for (int i = 0; i < 4; i++) (*this)[i] = r.f64();
}

// ---------------- POLYNOMIAL_4D ----------------

POLYNOMIAL_4D::POLYNOMIAL_4D(const std::vector<QUATERNION>& roots) : _roots(roots), _rootPowers(roots.size(), 1.0) {}

POLYNOMIAL_4D::POLYNOMIAL_4D(const std::vector<QUATERNION>& roots, const std::vector<Real>& powers)
    : _roots(roots), _rootPowers(powers) {
  if (powers.size() != roots.size()) throw Error(QJ_ERR_INVALID, "POLYNOMIAL_4D: one power per root");
}

void POLYNOMIAL_4D::addRoot(const QUATERNION& r, const Real& power) {
  _roots.push_back(r);
  _rootPowers.push_back(power);
}

POLYNOMIAL_4D& POLYNOMIAL_4D::operator*=(const Real& alpha) {
  for (QUATERNION& r : _roots)
    for (int i = 0; i < 4; i++) r[i] *= alpha;
  return *this;
}

POLYNOMIAL_4D& POLYNOMIAL_4D::operator+=(const VEC3F& v) {
  for (QUATERNION& r : _roots)
    for (int i = 0; i < 3; i++) r[i] += v[i];
  return *this;
}

POLYNOMIAL_4D& POLYNOMIAL_4D::operator-=(const VEC3F& v) {
  for (QUATERNION& r : _roots)
    for (int i = 0; i < 3; i++) r[i] -= v[i];
  return *this;
}

void POLYNOMIAL_4D::translateExceptFirst(const VEC3F& v) {
  for (size_t k = 1; k < _roots.size(); k++)
    for (int i = 0; i < 3; i++) _roots[k][i] += v[i];
}

void POLYNOMIAL_4D::clear() {
  _roots.clear();
  _rootPowers.clear();
  _powerScalar = 1.0;
}

void POLYNOMIAL_4D::write(FILE* file) const {
  Writer w{file};
  write_poly(w, *this);
}

void POLYNOMIAL_4D::read(FILE* file) {
  Reader r{file};
  *this = read_poly(r);
}

qj_poly POLYNOMIAL_4D::view() const {
  if (_rootPowers.size() != _roots.size()) throw Error(QJ_ERR_INVALID, "POLYNOMIAL_4D: one power per root");
  _flat.resize(4 * _roots.size());
  for (size_t i = 0; i < _roots.size(); i++)
    for (int k = 0; k < 4; k++) _flat[4 * i + k] = _roots[i][k];
  qj_poly p;
  p.n_roots = (int32_t)_roots.size();
  p.roots = _roots.empty() ? nullptr : _flat.data();
  p.powers = _roots.empty() ? nullptr : _rootPowers.data();
  p.power_scalar = _powerScalar;
  return p;
}

// ---------------- FIELD_3D ----------------

FIELD_3D::FIELD_3D(int xRes, int yRes, int zRes, const VEC3F& center, const VEC3F& lengths)
    : _xRes(xRes), _yRes(yRes), _zRes(zRes), _center(center), _lengths(lengths) {
  if (xRes < 0 || yRes < 0 || zRes < 0) throw Error(QJ_ERR_INVALID, "FIELD_3D: negative resolution");
  _data.assign(totalCells(), 0.0);
}

FIELD_3D::FIELD_3D(const Real* data, int xRes, int yRes, int zRes, const VEC3F& center, const VEC3F& lengths)
    : FIELD_3D(xRes, yRes, zRes, center, lengths) {
  std::copy(data, data + totalCells(), _data.begin());
}

FIELD_3D& FIELD_3D::operator=(Real value) {
  std::fill(_data.begin(), _data.end(), value);
  return *this;
}

qj_grid FIELD_3D::grid() const {
  qj_grid g;
  g.res[0] = _xRes;
  g.res[1] = _yRes;
  g.res[2] = _zRes;
  for (int i = 0; i < 3; i++) {
    g.center[i] = _center[i];
    g.lengths[i] = _lengths[i];
  }
  for (int i = 0; i < 4; i++) g.rotation[i] = _rotation[i];
  return g;
}

// FIELD_3D::cellCenter (FIELD_3D.cpp:818-841): ((c - L/2) + i dx + dx/2 - c)
// rotated about c, in this operation order.
VEC3F FIELD_3D::cellCenter(int x, int y, int z) const {
  const Real d[3] = {dx(), dy(), dz()};
  const int idx[3] = {x, y, z};
  Real f[3];
  for (int i = 0; i < 3; i++) {
    f[i] = _center[i] - 0.5 * _lengths[i];
    f[i] += idx[i] * d[i];
    f[i] += d[i] * 0.5;
    f[i] -= _center[i];
  }
  const Real w = _rotation[0], qx = _rotation[1], qy = _rotation[2], qz = _rotation[3];
  const Real m[9] = {1 - 2 * qy * qy - 2 * qz * qz, 2 * qx * qy - 2 * w * qz,     2 * qx * qz + 2 * w * qy,
                     2 * qx * qy + 2 * w * qz,     1 - 2 * qx * qx - 2 * qz * qz, 2 * qy * qz - 2 * w * qx,
                     2 * qx * qz - 2 * w * qy,     2 * qy * qz + 2 * w * qx,     1 - 2 * qx * qx - 2 * qy * qy};
  VEC3F out;
  for (int i = 0; i < 3; i++) {
    const Real r = m[3 * i] * f[0] + m[3 * i + 1] * f[1] + m[3 * i + 2] * f[2];
    out[i] = r + _center[i];
  }
  return out;
}

// FIELD_3D::operator()(VEC3F) (FIELD_3D.cpp:966-1029): clamped trilinear
// lookup with single-precision weights.
Real FIELD_3D::operator()(const VEC3F& position) const {
  const Real invDx = 1.0 / dx(), invDy = 1.0 / dy(), invDz = 1.0 / dz();
  Real c0 = _center[0] - 0.5 * _lengths[0], c1 = _center[1] - 0.5 * _lengths[1], c2 = _center[2] - 0.5 * _lengths[2];
  c0 += 0.5 * dx();
  c1 += 0.5 * dy();
  c2 += 0.5 * dz();
  Real p0 = position[0] - c0, p1 = position[1] - c1, p2 = position[2] - c2;
  p0 *= invDx;
  p1 *= invDy;
  p2 *= invDz;
  int x0 = (int)p0, y0 = (int)p1, z0 = (int)p2;
  int x1 = x0 + 1, y1 = y0 + 1, z1 = z0 + 1;
  auto clampi = [](int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); };
  x0 = clampi(x0, _xRes - 1);
  y0 = clampi(y0, _yRes - 1);
  z0 = clampi(z0, _zRes - 1);
  x1 = clampi(x1, _xRes - 1);
  y1 = clampi(y1, _yRes - 1);
  z1 = clampi(z1, _zRes - 1);
  const float s1 = (float)(p0 - x0), s0 = 1.0f - s1;
  const float t1 = (float)(p1 - y0), t0 = 1.0f - t1;
  const float u1 = (float)(p2 - z0), u0 = 1.0f - u1;
  const FIELD_3D& F = *this;
  return u0 * (s0 * (t0 * F(x0, y0, z0) + t1 * F(x0, y1, z0)) + s1 * (t0 * F(x1, y0, z0) + t1 * F(x1, y1, z0))) +
         u1 * (s0 * (t0 * F(x0, y0, z1) + t1 * F(x0, y1, z1)) + s1 * (t0 * F(x1, y0, z1) + t1 * F(x1, y1, z1)));
}

void FIELD_3D::write(FILE* file) const {
  Writer w{file};
  write_field(w, *this);
}

void FIELD_3D::read(FILE* file) {
  Reader r{file};
  const QUATERNION rot = _rotation;
  *this = read_field(r);
  _rotation = rot;
}

void FIELD_3D::write(const std::string& filename) const {
  Writer w;
  if (ends_with_gz(filename)) {
    w.g = gzopen(filename.c_str(), "wb1");
    if (!w.g) throw Error(QJ_ERR_INVALID, "FIELD_3D write: cannot open " + filename);
    try {
      write_field(w, *this);
    } catch (...) {
      gzclose(w.g);
      throw;
    }
    gzclose(w.g);
    return;
  }
  w.f = fopen(filename.c_str(), "wb");
  if (!w.f) throw Error(QJ_ERR_INVALID, "FIELD_3D write: cannot open " + filename);
  try {
    write_field(w, *this);
  } catch (...) {
    fclose(w.f);
    throw;
  }
  fclose(w.f);
}

void FIELD_3D::read(const std::string& filename) {
  Reader r;
  if (ends_with_gz(filename)) {
    r.g = gzopen(filename.c_str(), "rb");
    if (!r.g) throw Error(QJ_ERR_INVALID, "FIELD_3D read: cannot open " + filename);
  } else {
    r.f = fopen(filename.c_str(), "rb");
    if (!r.f) throw Error(QJ_ERR_INVALID, "FIELD_3D read: cannot open " + filename);
  }
  try {
    *this = read_field(r);
  } catch (...) {
    if (r.g) gzclose(r.g);
    if (r.f) fclose(r.f);
    throw;
  }
  if (r.g) gzclose(r.g);
  if (r.f) fclose(r.f);
}

// ---------------- TRIANGLE_MESH ----------------

namespace {
void fetch_mesh(qj_workspace* ws, int64_t nv, int64_t nt, std::vector<VEC3F>& v, std::vector<TRIANGLE>& t) {
  static_assert(sizeof(VEC3F) == 3 * sizeof(double), "VEC3F is 3 packed doubles");
  static_assert(sizeof(TRIANGLE) == 3 * sizeof(int32_t), "TRIANGLE is 3 packed ints");
  v.resize((size_t)nv);
  t.resize((size_t)nt);
  check(qj_mc_copy(ws, nv ? &v[0][0] : nullptr, nt ? (int32_t*)&t[0].v[0] : nullptr, nullptr), "qj_mc_copy");
}
}  // namespace

namespace {

// one FIELD_3D pass on the device: the cells go up, `op` runs, they come back
template <class Op>
void device_pass(FIELD_3D& field, Op op) {
  const size_t n = field.totalCells();
  if (!n) return;
  DeviceBuffer d(sizeof(double) * n);
  hip_check(hipMemcpy(d.p, field.data(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  op((double*)d.p);
  hip_check(hipMemcpy(field.data(), d.p, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
}

}  // namespace

void FIELD_3D::clampInfsToMean() {
  device_pass(*this, [&](double* d) {
    check(qj_field_clamp_infs_to_mean(shared_workspace(), d, (int64_t)totalCells(), nullptr, nullptr),
          "qj_field_clamp_infs_to_mean");
  });
}

void FIELD_3D::clampNansToMean() {
  device_pass(*this, [&](double* d) {
    check(qj_field_clamp_nans_to_mean(shared_workspace(), d, (int64_t)totalCells(), nullptr, nullptr),
          "qj_field_clamp_nans_to_mean");
  });
}

void FIELD_3D::setBorders(const Real& value) {
  const qj_grid g = grid();
  device_pass(*this, [&](double* d) { check(qj_field_set_borders(&g, value, d, nullptr), "qj_field_set_borders"); });
}

void FIELD_3D::scaleLengths(const Real& scale) {
  for (int i = 0; i < 3; i++) _lengths[i] *= scale;
}

FIELD_3D& FIELD_3D::operator*=(const Real& alpha) {
  for (Real& v : _data) v *= alpha;
  return *this;
}

Real FIELD_3D::sum() const {
  Real s = 0;
  for (Real v : _data) s += v;
  return s;
}

Real FIELD_3D::absSum() const {
  Real s = 0;
  for (Real v : _data) s += fabs(v);
  return s;
}

namespace {
void field_range(const FIELD_3D& f, Real* mn, Real* mx) {
  const size_t n = f.totalCells();
  if (!n) throw Error(QJ_ERR_INVALID, "FIELD_3D fieldMin/fieldMax: empty field");
  DeviceBuffer d(sizeof(double) * n);
  hip_check(hipMemcpy(d.p, f.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  check(qj_field_min_max(shared_workspace(), (const double*)d.p, (int64_t)n, mn, mx, nullptr), "qj_field_min_max");
}

FIELD_3D inverted(const FIELD_3D& dist, bool normalized, Real* abs_sum) {
  FIELD_3D out(dist);
  const size_t n = dist.totalCells();
  if (abs_sum) *abs_sum = 0;
  if (!n) return out;
  DeviceBuffer d(sizeof(double) * n), o(sizeof(double) * n);
  hip_check(hipMemcpy(d.p, dist.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  const qj_grid g = dist.grid();
  check(qj_field_inverted_distance(shared_workspace(), &g, (const double*)d.p, normalized ? 1 : 0, (double*)o.p,
                                   abs_sum, nullptr),
        "qj_field_inverted_distance");
  hip_check(hipMemcpy(out.data(), o.p, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
  return out;
}
}  // namespace

Real FIELD_3D::fieldMax() const {
  Real mn, mx;
  field_range(*this, &mn, &mx);
  return mx;
}

Real FIELD_3D::fieldMin() const {
  Real mn, mx;
  field_range(*this, &mn, &mx);
  return mn;
}

FIELD_3D FIELD_3D::invertedDistance() const { return inverted(*this, false, nullptr); }

FIELD_3D FIELD_3D::invertedDistanceNormalized(Real* absSum) const { return inverted(*this, true, absSum); }

TRIANGLE_MESH::TRIANGLE_MESH(const FIELD_3D& field, Real isosurface) {
  qj_workspace* ws = shared_workspace();
  DeviceBuffer d(sizeof(double) * field.totalCells());
  if (field.totalCells())
    hip_check(hipMemcpy(d.p, field.dataConst(), sizeof(double) * field.totalCells(), hipMemcpyHostToDevice),
              "hipMemcpy field");
  const qj_grid g = field.grid();
  int64_t nv = 0, nt = 0;
  check(qj_mc_run(ws, (const double*)d.p, &g, isosurface, &nv, &nt, nullptr), "qj_mc_run");
  fetch_mesh(ws, nv, nt, _vertices, _triangles);
}

namespace {

qj_grid mesher_grid(const VEC3F& center, const VEC3F& lengths, const VEC3I& res, const QUATERNION& rotation) {
  qj_grid g;
  for (int i = 0; i < 3; i++) {
    g.res[i] = res[i];
    g.center[i] = center[i];
    g.lengths[i] = lengths[i];
  }
  for (int i = 0; i < 4; i++) g.rotation[i] = rotation[i];
  return g;
}

// readEdgeCacheHuge (TRIANGLE_MESH.cpp:431-481): int count; per vertex the
// two grid points of its edge (3 + 3 ints); int count; per active cell the
// flag and the cell (1 + 3 ints). Returns false when the file does not exist.
bool read_edge_cache(const std::string& name, std::vector<int32_t>& edges, std::vector<int32_t>& cells) {
  FILE* f = fopen(name.c_str(), "rb");
  if (!f) return false;
  Reader r;
  r.f = f;
  try {
    const int n = r.i32();
    if (n < 0) throw Error(QJ_ERR_INVALID, "edge cache: negative triplet count");
    edges.resize(4 * (size_t)n);
    for (int i = 0; i < n; i++) {
      int a[6];
      r.get(a, sizeof a);
      const int dx = a[3] - a[0], dy = a[4] - a[1], dz = a[5] - a[2];
      int axis = -1;
      if (dx == 1 && dy == 0 && dz == 0) axis = 0;
      if (dx == 0 && dy == 1 && dz == 0) axis = 1;
      if (dx == 0 && dy == 0 && dz == 1) axis = 2;
      if (axis < 0) throw Error(QJ_ERR_INVALID, "edge cache: a triplet is not a +x/+y/+z grid edge");
      edges[4 * i] = a[0];
      edges[4 * i + 1] = a[1];
      edges[4 * i + 2] = a[2];
      edges[4 * i + 3] = axis;
    }
    const int m = r.i32();
    if (m < 0) throw Error(QJ_ERR_INVALID, "edge cache: negative flag count");
    cells.resize(4 * (size_t)m);
    for (int i = 0; i < m; i++) {
      int a[4];
      r.get(a, sizeof a);
      cells[4 * i] = a[1];
      cells[4 * i + 1] = a[2];
      cells[4 * i + 2] = a[3];
      cells[4 * i + 3] = a[0];
    }
  } catch (...) {
    fclose(f);
    throw;
  }
  fclose(f);
  return true;
}

// writeEdgeCacheHuge (TRIANGLE_MESH.cpp:531-583)
void write_edge_cache(const std::string& name, const std::vector<int32_t>& edges, const std::vector<int32_t>& cells) {
  Writer w;
  w.f = fopen(name.c_str(), "wb");
  if (!w.f) throw Error(QJ_ERR_INVALID, "edge cache: cannot open " + name);
  try {
    const int n = (int)(edges.size() / 4), m = (int)(cells.size() / 4);
    w.i32(n);
    for (int i = 0; i < n; i++) {
      const int32_t* e = &edges[4 * (size_t)i];
      const int a[6] = {e[0], e[1], e[2], e[0] + (e[3] == 0), e[1] + (e[3] == 1), e[2] + (e[3] == 2)};
      w.put(a, sizeof a);
    }
    w.i32(m);
    for (int i = 0; i < m; i++) {
      const int32_t* c = &cells[4 * (size_t)i];
      const int a[4] = {c[3], c[0], c[1], c[2]};
      w.put(a, sizeof a);
    }
  } catch (...) {
    fclose(w.f);
    throw;
  }
  fclose(w.f);
}

}  // namespace

TRIANGLE_MESH::TRIANGLE_MESH(const VEC3F& center, const VEC3F& lengths, const VEC3I& res, const POLYNOMIAL_4D& top,
                             const POLYNOMIAL_4D& bottom, Real expScaling, int maxIterations, Real slice,
                             Real isosurface, const QUATERNION& rotation, const std::string& cacheFilename,
                             qj_precision precision, int chunkSlices) {
  qj_workspace* ws = shared_workspace();
  const qj_grid g = mesher_grid(center, lengths, res, rotation);
  const qj_poly t = top.view(), b = bottom.view();
  qj_rational* f = nullptr;
  check(qj_rational_create(&t, &b, &f), "qj_rational_create");
  // the mesher's map: escape 2000 and the black-hole stop (TRIANGLE_MESH.cpp:97, 1789-1791)
  const qj_map_params p{maxIterations, 2000.0, expScaling, slice, QJ_MAP_BLACK_HOLE};
  int64_t nv = 0, nt = 0;
  std::vector<int32_t> edges, cells;
  qj_status st;
  try {
    if (!cacheFilename.empty() && read_edge_cache(cacheFilename, edges, cells)) {
      st = qj_nlmc_resume(ws, &g, f, &p, isosurface, (int64_t)(edges.size() / 4), edges.data(),
                          (int64_t)(cells.size() / 4), cells.data(), &nv, &nt, nullptr);
    } else {
      st = qj_nlmc_run(ws, &g, f, &p, isosurface, precision, 0.05, chunkSlices, &nv, &nt, nullptr);
      if (st == QJ_OK && !cacheFilename.empty()) {
        int64_t ne = 0, nc = 0;
        check(qj_nlmc_cache(ws, &ne, nullptr, &nc, nullptr, nullptr), "qj_nlmc_cache");
        edges.resize(4 * (size_t)ne);
        cells.resize(4 * (size_t)nc);
        check(qj_nlmc_cache(ws, nullptr, edges.data(), nullptr, cells.data(), nullptr), "qj_nlmc_cache");
        write_edge_cache(cacheFilename, edges, cells);
      }
    }
  } catch (...) {
    qj_rational_destroy(f);
    throw;
  }
  qj_rational_destroy(f);
  check(st, "non-linear mesher");
  fetch_mesh(ws, nv, nt, _vertices, _triangles);
}

TRIANGLE_MESH::TRIANGLE_MESH(const FIELD_3D& field, const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom,
                             Real expScaling, int maxIterations, Real slice, Real isosurface) {
  qj_workspace* ws = shared_workspace();
  DeviceBuffer d(sizeof(double) * field.totalCells());
  if (field.totalCells())
    hip_check(hipMemcpy(d.p, field.dataConst(), sizeof(double) * field.totalCells(), hipMemcpyHostToDevice),
              "hipMemcpy field");
  const qj_grid g = field.grid();
  const qj_poly t = top.view(), b = bottom.view();
  qj_rational* f = nullptr;
  check(qj_rational_create(&t, &b, &f), "qj_rational_create");
  // _escapeRadius = 2000 (TRIANGLE_MESH.cpp:76); the black-hole stop of nonlinearValue
  const qj_map_params p{maxIterations, 2000.0, expScaling, slice, QJ_MAP_BLACK_HOLE};
  int64_t nv = 0, nt = 0;
  const qj_status st = qj_nlmc_field(ws, (const double*)d.p, &g, f, &p, isosurface, &nv, &nt, nullptr);
  qj_rational_destroy(f);
  check(st, "non-linear mesher on a field");
  fetch_mesh(ws, nv, nt, _vertices, _triangles);
}

TRIANGLE_MESH::TRIANGLE_MESH(const VEC3F& center, const VEC3F& lengths, const VEC3I& res, const POLYNOMIAL_4D& top,
                             const POLYNOMIAL_4D& bottom, Real expScaling, int maxIterations, Real slice,
                             Real isosurface, const QUATERNION& rotation) {
  qj_workspace* ws = shared_workspace();
  const qj_grid g = mesher_grid(center, lengths, res, rotation);
  const qj_poly t = top.view(), b = bottom.view();
  qj_rational* f = nullptr;
  check(qj_rational_create(&t, &b, &f), "qj_rational_create");
  const qj_map_params p{maxIterations, 2000.0, expScaling, slice, QJ_MAP_BLACK_HOLE | QJ_MAP_LAGRANGE};
  int64_t nv = 0, nt = 0;
  const qj_status st = qj_nlmc_run(ws, &g, f, &p, isosurface, QJ_FP64, 0.05, 64, &nv, &nt, nullptr);
  qj_rational_destroy(f);
  check(st, "Lagrange mesher");
  fetch_mesh(ws, nv, nt, _vertices, _triangles);
}

// TRIANGLE_MESH::writeOBJ (TRIANGLE_MESH.cpp:2762-2800): "v %.16f %.16f %.16f"
// per vertex, then "f i j k" with 1-based indices (no normals are computed
// by the marching-cubes constructors).
bool TRIANGLE_MESH::writeOBJ(const std::string& filename) const {
  FILE* file = fopen(filename.c_str(), "w");
  if (!file) return false;
  for (const VEC3F& v : _vertices) fprintf(file, "v %.16f %.16f %.16f\n", v[0], v[1], v[2]);
  for (const TRIANGLE& t : _triangles) fprintf(file, "f %i %i %i\n", t.v[0] + 1, t.v[1] + 1, t.v[2] + 1);
  fclose(file);
  return true;
}

// ---------------- OPTIMIZE_3D ----------------

OPTIMIZE_3D::OPTIMIZE_3D() {}

OPTIMIZE_3D::OPTIMIZE_3D(int xRes, int yRes, int zRes, const VEC3F& center, const VEC3F& lengths)
    : _xRes(xRes), _yRes(yRes), _zRes(zRes), _center(center), _lengths(lengths),
      _fractal(xRes, yRes, zRes, center, lengths) {}

OPTIMIZE_3D::~OPTIMIZE_3D() {
  if (_dField) (void)hipFree(_dField);
  if (_dWeight) (void)hipFree(_dWeight);
  if (_rational) qj_rational_destroy(_rational);
  if (_ws) qj_workspace_destroy(_ws);
}

void OPTIMIZE_3D::sync_rational(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom) {
  const qj_poly t = top.view(), b = bottom.view();
  if (!_ws) check(qj_workspace_create(&_ws), "qj_workspace_create");
  if (_rational) {
    int32_t nt = 0, nb = 0;
    check(qj_rational_counts(_rational, &nt, &nb), "qj_rational_counts");
    if (nt == t.n_roots && nb == b.n_roots) {
      check(qj_rational_update(_rational, &t, &b, nullptr), "qj_rational_update");
      return;
    }
    qj_rational_destroy(_rational);
    _rational = nullptr;
  }
  check(qj_rational_create(&t, &b, &_rational), "qj_rational_create");
}

void OPTIMIZE_3D::ensure_device(size_t cells) {
  if (cells <= _dCells && _dField) return;
  if (_dField) hip_check(hipFree(_dField), "hipFree");
  _dField = nullptr;
  _dCells = 0;
  if (cells) hip_check(hipMalloc((void**)&_dField, sizeof(double) * cells), "hipMalloc field");
  _dCells = cells;
}

void OPTIMIZE_3D::computeLogScaledPowerRationalMap(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom,
                                                   FIELD_3D& fractal, bool verbose) {
  (void)verbose;
  sync_rational(top, bottom);
  const size_t n = fractal.totalCells();
  if (!n) return;
  ensure_device(n);
  const qj_grid g = fractal.grid();
  const qj_map_params p{_maxIterations, 2000.0, _expScaling, _quaternionSlice, QJ_MAP_BLACK_HOLE};
  check(qj_field_eval(_ws, &g, _rational, &p, _precision, _band, _dField, nullptr), "qj_field_eval");
  hip_check(hipMemcpy(fractal.data(), _dField, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
  _apValid = false;
}

void OPTIMIZE_3D::computeRationalMapAP(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, FIELD_3D& fractal,
                                       bool verbose) {
  (void)verbose;
  sync_rational(top, bottom);
  const size_t n = fractal.totalCells();
  if (!n) return;
  ensure_device(n);
  // cells at the origin are skipped and keep their previous value (OPTIMIZE_3D.cpp:1432-1435)
  hip_check(hipMemcpy(_dField, fractal.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  const qj_grid g = fractal.grid();
  const qj_map_params p{_maxIterations, 20.0, _expScaling, _quaternionSlice, QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE};
  check(qj_field_eval(_ws, &g, _rational, &p, QJ_FP64, 0.0, _dField, nullptr), "qj_field_eval");
  hip_check(hipMemcpy(fractal.data(), _dField, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
  _apValid = false;
}

void OPTIMIZE_3D::map_into(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, FIELD_3D& fractal,
                           const qj_map_params& p, bool keepSkipped) {
  sync_rational(top, bottom);
  const size_t n = fractal.totalCells();
  if (!n) return;
  ensure_device(n);
  if (keepSkipped)
    hip_check(hipMemcpy(_dField, fractal.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  const qj_grid g = fractal.grid();
  check(qj_field_eval(_ws, &g, _rational, &p, QJ_FP64, 0.0, _dField, nullptr), "qj_field_eval");
  hip_check(hipMemcpy(fractal.data(), _dField, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
  _apValid = false;
}

void OPTIMIZE_3D::computeLogScaledPowerRationalMapCached(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom,
                                                         FIELD_3D& fractal, bool verbose) {
  (void)verbose;
  // cells at the origin are skipped and keep their previous value (OPTIMIZE_3D.cpp:1624-1625)
  const qj_map_params p{_maxIterations, 20.0, _expScaling, _quaternionSlice, QJ_MAP_SKIP_ORIGIN};
  map_into(top, bottom, fractal, p, true);
}

void OPTIMIZE_3D::computeLogPowerRationalMap(bool verbose) {
  (void)verbose;
  // evaluatePowerFactored: the root powers without the power scalar
  POLYNOMIAL_4D top(_top), bottom(_bottom);
  top.powerScalar() = 1.0;
  bottom.powerScalar() = 1.0;
  const qj_map_params p{_maxIterations, 20.0, _expScaling, _quaternionSlice, 0u};
  map_into(top, bottom, _fractal, p, false);
}

void OPTIMIZE_3D::scores_on_device(const FIELD_3D& logFractal, const FIELD_3D& weight,
                                   const std::vector<Real>& thresholds, std::vector<Real>* scores, Real* maxScore) {
  const size_t n = logFractal.totalCells(), wn = weight.totalCells();
  if (!n) throw Error(QJ_ERR_INVALID, "OPTIMIZE_3D score: the fractal is empty");
  if (!wn) throw Error(QJ_ERR_INVALID, "OPTIMIZE_3D score: the weight field is empty");
  if (!_ws) check(qj_workspace_create(&_ws), "qj_workspace_create");
  ensure_device(n);
  hip_check(hipMemcpy(_dField, logFractal.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice),
            "hipMemcpy field");
  if (wn > _dWeightCells || !_dWeight) {
    if (_dWeight) hip_check(hipFree(_dWeight), "hipFree");
    _dWeight = nullptr;
    hip_check(hipMalloc((void**)&_dWeight, sizeof(double) * wn), "hipMalloc weight");
    _dWeightCells = wn;
  }
  hip_check(hipMemcpy(_dWeight, weight.dataConst(), sizeof(double) * wn, hipMemcpyHostToDevice), "hipMemcpy weight");
  const qj_grid g = logFractal.grid(), wg = weight.grid();
  if (scores) scores->assign(thresholds.size(), 0.0);
  check(qj_weight_scores(_ws, &g, _dField, &wg, _dWeight, _tanhAlpha, thresholds.data(), (int32_t)thresholds.size(),
                         scores ? scores->data() : nullptr, maxScore, nullptr),
        "qj_weight_scores");
}

Real OPTIMIZE_3D::weightScore(const FIELD_3D& logFractal, const FIELD_3D& weightField) {
  std::vector<Real> s;
  scores_on_device(logFractal, weightField, std::vector<Real>{0.0}, &s, nullptr);
  return s[0];
}

Real OPTIMIZE_3D::inverseDistanceScore(const FIELD_3D& logFractal, const FIELD_3D& inverseDistanceField,
                                       const Real& inverseSum, const Real& threshold) {
  (void)inverseSum;  // the reference leaves its normalisation commented out (OPTIMIZE_3D.cpp:4560)
  std::vector<Real> s;
  scores_on_device(logFractal, inverseDistanceField, std::vector<Real>{threshold}, &s, nullptr);
  return s[0];
}

std::vector<Real> OPTIMIZE_3D::inverseDistanceScores(const FIELD_3D& logFractal, const FIELD_3D& inverseDistanceField,
                                                     const std::vector<Real>& thresholds) {
  std::vector<Real> s;
  scores_on_device(logFractal, inverseDistanceField, thresholds, &s, nullptr);
  return s;
}

Real OPTIMIZE_3D::computeLogScaledPowerRationalScore() {
  computeLogScaledPowerRationalMapCached(_top, _bottom, _fractal);
  return weightScore(_fractal, _weightField);
}

Real OPTIMIZE_3D::computeMaximumScaledPowerRationalScore() {
  Real m = 0;
  scores_on_device(_fractal, _weightField, std::vector<Real>(), nullptr, &m);
  return m;
}

Real OPTIMIZE_3D::estimateConformalRadius() {
  const VEC3F center = _fractal.center(), lengths = _fractal.lengths();
  const int fractalRes = _fractal.xRes();
  _fractal = FIELD_3D(fractalRes, fractalRes, fractalRes, center, lengths);
  const int oldMaxIterations = _maxIterations;
  _expScaling = 1;
  _maxIterations = 1;
  computeLogPowerRationalMap(true);
  _fractal.clampInfsToMean();
  const FIELD_3D inverseDistanceField = _distanceField.invertedDistance();
  Real minScore = _fractal.fieldMin(), maxScore = _fractal.fieldMax();
  Real minThreshold = minScore;
  Real minScoreFound = FLT_MAX;
  for (int i = 0; i < 4; i++) {
    const Real increment = (maxScore - minScore) / 100;
    minScoreFound = FLT_MAX;
    // the reference's scan, all thresholds of a pass scored in one GPU pass
    std::vector<Real> xs;
    for (Real x = minScore; x < maxScore; x += increment) {
      if (xs.size() >= 4096 || x + increment == x)
        throw Error(QJ_ERR_NUMERIC, "estimateConformalRadius: the threshold scan does not advance");
      xs.push_back(x);
    }
    const std::vector<Real> sc = xs.empty() ? std::vector<Real>() : inverseDistanceScores(_fractal,
                                                                                          inverseDistanceField, xs);
    for (size_t k = 0; k < xs.size(); k++)
      if (sc[k] < minScoreFound) {
        minScoreFound = sc[k];
        minThreshold = xs[k];
      }
    minScore = minThreshold - 5 * increment;
    maxScore = minThreshold + 5 * increment;
  }
  _expScaling = exp(-minThreshold);
  _maxIterations = oldMaxIterations;
  return minScoreFound;
}

void OPTIMIZE_3D::setScoreFields(const FIELD_3D& distanceField, const FIELD_3D& curvatureField) {
  if (distanceField.xRes() != curvatureField.xRes() || distanceField.yRes() != curvatureField.yRes() ||
      distanceField.zRes() != curvatureField.zRes())
    throw Error(QJ_ERR_INVALID, "setScoreFields: the distance and curvature fields differ in resolution");
  _xRes = distanceField.xRes();
  _yRes = distanceField.yRes();
  _zRes = distanceField.zRes();
  _center = distanceField.center();
  _lengths = distanceField.lengths();
  FIELD_3D d(distanceField.dataConst(), _xRes, _yRes, _zRes, _center, _lengths);
  FIELD_3D c(curvatureField.dataConst(), _xRes, _yRes, _zRes, _center, _lengths);
  _distanceField = std::move(d);
  _curvatureField = std::move(c);
  _fractal = FIELD_3D(_xRes, _yRes, _zRes, _center, _lengths);
  _apValid = false;
}

void OPTIMIZE_3D::scaleEverything(const Real& scale) {
  for (int i = 0; i < 3; i++) {
    _lengths[i] *= scale;
    _center[i] *= scale;
  }
  for (FIELD_3D* f : {&_distanceField, &_inverseDistanceField, &_weightField, &_curvatureField, &_fractal}) {
    f->scaleLengths(scale);
    for (int i = 0; i < 3; i++) f->center()[i] *= scale;
  }
  _distanceField *= scale;
  _top *= scale;
  _bottom *= scale;
  _apValid = false;
}

void OPTIMIZE_3D::translateEverything(const VEC3F& translation) {
  for (int i = 0; i < 3; i++) _center[i] += translation[i];
  for (FIELD_3D* f : {&_distanceField, &_inverseDistanceField, &_weightField, &_curvatureField, &_fractal})
    for (int i = 0; i < 3; i++) f->center()[i] += translation[i];
  _top += translation;
  _bottom += translation;
  _apValid = false;
}

void OPTIMIZE_3D::translateEverythingExceptFirst(const VEC3F& translation) {
  for (int i = 0; i < 3; i++) _center[i] += translation[i];
  for (FIELD_3D* f : {&_distanceField, &_inverseDistanceField, &_weightField, &_curvatureField, &_fractal})
    for (int i = 0; i < 3; i++) f->center()[i] += translation[i];
  _top.translateExceptFirst(translation);
  _bottom.translateExceptFirst(translation);
  _apValid = false;
}

// One fused GPU pass gives the AP field, the score and every gradient for the
// current state; the accessors reuse it while the state (roots, powers,
// scalars, c, alpha, threshold, weight field) is unchanged.
void OPTIMIZE_3D::run_ap() {
  if (_maxIterations != 1)
    throw Error(QJ_ERR_INVALID,
                "OPTIMIZE_3D AP score/gradients: maxIterations must be 1 (the gradients are one-iteration "
                "closed forms; the optimizer sets it to 1, toothOptimizer.cpp:402)");
  if (_weightField.totalCells() == 0) throw Error(QJ_ERR_INVALID, "OPTIMIZE_3D: weight field is empty");
  std::vector<double> key;
  auto push_poly = [&key](const POLYNOMIAL_4D& p) {
    key.push_back(p.totalRoots());
    for (const QUATERNION& q : p.roots())
      for (int i = 0; i < 4; i++) key.push_back(q[i]);
    for (Real e : p.powers()) key.push_back(e);
    key.push_back(p.powerScalar());
  };
  push_poly(_top);
  push_poly(_bottom);
  key.push_back(_expScaling);
  key.push_back(_quaternionSlice);
  key.push_back(_tanhAlpha);
  key.push_back(_tanhThreshold);
  key.push_back(_fractal.xRes());
  key.push_back(_fractal.yRes());
  key.push_back(_fractal.zRes());
  // weight field identity: FNV-1a of its bytes
  uint64_t h = 1469598103934665603ull;
  const unsigned char* wb = (const unsigned char*)_weightField.dataConst();
  for (size_t i = 0; i < sizeof(double) * _weightField.totalCells(); i++) h = (h ^ wb[i]) * 1099511628211ull;
  double hd;
  memcpy(&hd, &h, sizeof hd);
  key.push_back(hd);
  key.push_back(_weightField.xRes());
  key.push_back(_weightField.yRes());
  key.push_back(_weightField.zRes());
  if (_apValid && key.size() == _apKey.size() && memcmp(key.data(), _apKey.data(), key.size() * sizeof(double)) == 0)
    return;

  sync_rational(_top, _bottom);
  const size_t n = _fractal.totalCells();
  ensure_device(n);
  hip_check(hipMemcpy(_dField, _fractal.dataConst(), sizeof(double) * n, hipMemcpyHostToDevice), "hipMemcpy field");
  const size_t wn = _weightField.totalCells();
  if (wn > _dWeightCells || !_dWeight) {
    if (_dWeight) hip_check(hipFree(_dWeight), "hipFree");
    _dWeight = nullptr;
    hip_check(hipMalloc((void**)&_dWeight, sizeof(double) * wn), "hipMalloc weight");
    _dWeightCells = wn;
  }
  hip_check(hipMemcpy(_dWeight, _weightField.dataConst(), sizeof(double) * wn, hipMemcpyHostToDevice),
            "hipMemcpy weight");
  const qj_grid g = _fractal.grid(), wg = _weightField.grid();
  const qj_map_params p{1, 20.0, _expScaling, _quaternionSlice, QJ_MAP_SKIP_ORIGIN | QJ_MAP_ADAPTIVE};
  _gTop.assign(std::max(1, _top.totalRoots()), 0.0);
  _gBottom.assign(std::max(1, _bottom.totalRoots()), 0.0);
  check(qj_ap_score_gradients(_ws, &g, _rational, &p, &wg, _dWeight, _tanhAlpha, _tanhThreshold, _dField,
                              _gTop.data(), _gBottom.data(), &_ap, nullptr),
        "qj_ap_score_gradients");
  hip_check(hipMemcpy(_fractal.data(), _dField, sizeof(double) * n, hipMemcpyDeviceToHost), "hipMemcpy field");
  _apKey = key;
  _apValid = true;
}

Real OPTIMIZE_3D::computeScoreAP() {
  run_ap();
  return _ap.score;
}
Real OPTIMIZE_3D::computeExpGradientAP() {
  run_ap();
  return _ap.grad_exp;
}
Real OPTIMIZE_3D::computeTopGradientAP(int whichRoot) {
  run_ap();
  if (whichRoot < 0 || whichRoot >= _top.totalRoots()) throw Error(QJ_ERR_INVALID, "computeTopGradientAP: root");
  return _gTop[whichRoot];
}
Real OPTIMIZE_3D::computeBottomGradientAP(int whichRoot) {
  run_ap();
  if (whichRoot < 0 || whichRoot >= _bottom.totalRoots())
    throw Error(QJ_ERR_INVALID, "computeBottomGradientAP: root");
  return _gBottom[whichRoot];
}
Real OPTIMIZE_3D::computeTopBulkScalarGradientAP() {
  run_ap();
  return _ap.grad_top_scalar;
}
Real OPTIMIZE_3D::computeBottomBulkScalarGradientAP() {
  run_ap();
  return _ap.grad_bottom_scalar;
}
int OPTIMIZE_3D::adaptiveCells() {
  run_ap();
  return (int)_ap.adaptive_cells;
}

void OPTIMIZE_3D::write(const std::string& filename) const {
  Writer w;
  w.f = fopen(filename.c_str(), "wb");
  if (!w.f) throw Error(QJ_ERR_INVALID, "OPTIMIZE_3D write: cannot open " + filename);
  try {
    w.i32(_xRes);
    w.i32(_yRes);
    w.i32(_zRes);
    write_vec3(w, _center);
    write_vec3(w, _lengths);
    write_poly(w, _top);
    write_poly(w, _bottom);
    w.f64(_expScaling);
    w.i32(_historySize);
    write_field(w, _fractal);
    write_field(w, _distanceField);
    write_field(w, _curvatureField);
  } catch (...) {
    fclose(w.f);
    throw;
  }
  fclose(w.f);
}

void OPTIMIZE_3D::read(const std::string& filename) {
  readState(filename);
  // the reference's post-read set-up (OPTIMIZE_3D.cpp:3348-3388)
  setScoreFields(_distanceField, _distanceField);
  if (_inverseDistanceField.totalCells() == 0) {
    _inverseDistanceField = _distanceField.invertedDistanceNormalized();
    _inverseFieldSum = _inverseDistanceField.absSum();
  }
  computeWeightField();
}

void OPTIMIZE_3D::readState(const std::string& filename) {
  Reader r;
  r.f = fopen(filename.c_str(), "rb");
  if (!r.f) throw Error(QJ_ERR_INVALID, "OPTIMIZE_3D read: cannot open " + filename);
  try {
    _xRes = r.i32();
    _yRes = r.i32();
    _zRes = r.i32();
    _center = read_vec3(r);
    _lengths = read_vec3(r);
    _top = read_poly(r);
    _bottom = read_poly(r);
    _expScaling = r.f64();
    _historySize = r.i32();
    _fractal = read_field(r);
    _distanceField = read_field(r);
    _curvatureField = read_field(r);
  } catch (...) {
    fclose(r.f);
    throw;
  }
  fclose(r.f);
  _apValid = false;
}

}  // namespace quijibo
