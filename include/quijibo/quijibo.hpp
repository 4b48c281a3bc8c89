// C++ host layer that keeps QUIJIBO's class API on top of the C ABI (qj_c_api.h).
//
// Class and member names follow the reference so a caller of the reference can
// switch: POLYNOMIAL_4D (src/algebra/POLYNOMIAL_4D.h), FIELD_3D
// (src/geometry/FIELD_3D.h), TRIANGLE_MESH (src/geometry/TRIANGLE_MESH.h),
// and OPTIMIZE_3D (src/solvers/OPTIMIZE_3D.h, the map / score / gradient
// members and the .o3d file).
// Real is double (the reference's USING_DOUBLE_NLS build). All field work runs
// on the GPU through qj_* calls; data that the reference keeps in host memory
// (FIELD_3D's cells, meshes) is copied back to host memory here.
#pragma once

#include <stdio.h>

#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../qj_c_api.h"

namespace quijibo {

typedef double Real;

// Error raised where the reference calls exit(0).
struct Error : std::runtime_error {
  qj_status status;
  Error(qj_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};
void check(qj_status s, const char* where);

struct VEC3F {
  Real v[3];
  VEC3F() : v{0, 0, 0} {}
  VEC3F(Real a, Real b, Real c) : v{a, b, c} {}
  Real& operator[](int i) { return v[i]; }
  Real operator[](int i) const { return v[i]; }
  void write(FILE* f) const;  // 3 doubles (VEC3.h:238-256)
  void read(FILE* f);
};

struct VEC3I {  // integer triple (VEC3.h TVEC3<int>)
  int v[3];
  VEC3I() : v{0, 0, 0} {}
  VEC3I(int a, int b, int c) : v{a, b, c} {}
  int& operator[](int i) { return v[i]; }
  int operator[](int i) const { return v[i]; }
};

// Storage of a quaternion, slot order (w, x, y, z) (QUATERNION.h:31-155).
struct QUATERNION {
  Real _w = 0, _x = 0, _y = 0, _z = 0;
  QUATERNION() {}
  QUATERNION(Real w, Real x, Real y, Real z) : _w(w), _x(x), _y(y), _z(z) {}
  Real& operator[](int i) { return (&_w)[i]; }
  Real operator[](int i) const { return (&_w)[i]; }
  Real w() const { return _w; }
  Real x() const { return _x; }
  Real y() const { return _y; }
  Real z() const { return _z; }
  void write(FILE* f) const;  // 4 doubles (QUATERNION.cpp:286-306)
  void read(FILE* f);
};

class POLYNOMIAL_4D {
 public:
  POLYNOMIAL_4D() {}
  explicit POLYNOMIAL_4D(const std::vector<QUATERNION>& roots);
  POLYNOMIAL_4D(const std::vector<QUATERNION>& roots, const std::vector<Real>& powers);

  int totalRoots() const { return (int)_roots.size(); }
  const std::vector<QUATERNION>& roots() const { return _roots; }
  std::vector<QUATERNION>& rootsMutable() { return _roots; }
  const std::vector<Real>& powers() const { return _rootPowers; }
  std::vector<Real>& powersMutable() { return _rootPowers; }
  Real& powerScalar() { return _powerScalar; }
  Real powerScalar() const { return _powerScalar; }

  // every root times alpha, all four components (POLYNOMIAL_4D.cpp:1184-1190)
  POLYNOMIAL_4D& operator*=(const Real& alpha);
  // roots moved by (v, 0) (POLYNOMIAL_4D.cpp:1195-1216); translateExceptFirst
  // leaves root 0 where it is (:1220-1226)
  POLYNOMIAL_4D& operator+=(const VEC3F& v);
  POLYNOMIAL_4D& operator-=(const VEC3F& v);
  void translateExceptFirst(const VEC3F& v);

  void addRoot(const QUATERNION& r, const Real& power = 1.0);
  void modifyRoot(int which, const QUATERNION& r) { _roots.at(which) = r; }
  void changePower(int which, const Real& p) { _rootPowers.at(which) = p; }
  void clear();

  // POLYNOMIAL_4D::write/read layout (POLYNOMIAL_4D.cpp:650-728); the expanded
  // coefficient arrays are written empty and skipped on read.
  void write(FILE* file) const;
  void read(FILE* file);

  // flattened arrays for the C ABI
  qj_poly view() const;

 private:
  std::vector<QUATERNION> _roots;
  std::vector<Real> _rootPowers;
  Real _powerScalar = 1.0;
  mutable std::vector<double> _flat;
};

class FIELD_3D {
 public:
  FIELD_3D() {}
  FIELD_3D(int xRes, int yRes, int zRes, const VEC3F& center = VEC3F(0, 0, 0),
           const VEC3F& lengths = VEC3F(1, 1, 1));
  FIELD_3D(const Real* data, int xRes, int yRes, int zRes, const VEC3F& center, const VEC3F& lengths);
  FIELD_3D& operator=(Real value);  // fill

  Real& operator()(int x, int y, int z) { return _data[x + y * (size_t)_xRes + z * (size_t)_xRes * _yRes]; }
  Real operator()(int x, int y, int z) const { return _data[x + y * (size_t)_xRes + z * (size_t)_xRes * _yRes]; }
  Real& operator[](size_t i) { return _data[i]; }
  Real operator[](size_t i) const { return _data[i]; }
  Real operator()(const VEC3F& position) const;  // trilinear lookup (FIELD_3D.cpp:966-1029)

  int xRes() const { return _xRes; }
  int yRes() const { return _yRes; }
  int zRes() const { return _zRes; }
  size_t totalCells() const { return (size_t)_xRes * _yRes * _zRes; }
  Real dx() const { return _lengths[0] / _xRes; }
  Real dy() const { return _lengths[1] / _yRes; }
  Real dz() const { return _lengths[2] / _zRes; }
  const VEC3F& center() const { return _center; }
  const VEC3F& lengths() const { return _lengths; }
  VEC3F& center() { return _center; }
  // lengths times scale (FIELD_3D.cpp:5302-5311; the spacings follow)
  void scaleLengths(const Real& scale);
  QUATERNION& rotation() { return _rotation; }
  const QUATERNION& rotation() const { return _rotation; }
  Real* data() { return _data.data(); }
  const Real* dataConst() const { return _data.data(); }
  VEC3F cellCenter(int x, int y, int z) const;  // FIELD_3D.cpp:818-841

  // FIELD_3D.cpp:4371-4395 / 4400-4450: int xRes, yRes, zRes; center and
  // lengths as 3 doubles each; then the cells as doubles, x fastest. The
  // rotation is not stored (the reference's format has no field for it).
  void write(FILE* file) const;
  void read(FILE* file);
  // FIELD_3D.cpp:280-305, 378-390: a name ending in "gz" is read and written
  // through zlib with the same layout
  void write(const std::string& filename) const;
  void read(const std::string& filename);

  qj_grid grid() const;

  // Clean-up before meshing, run on the GPU (FIELD_3D.cpp:5182-5234,
  // 5357-5382): infinite (NaN) cells take the mean of the other cells;
  // setBorders sets the six faces to a constant.
  void clampInfsToMean();
  void clampNansToMean();
  void setBorders(const Real& value);

  FIELD_3D& operator*=(const Real& alpha);  // every cell times alpha
  // serial sums in index order (FIELD_3D.cpp:2820-2839), host side
  Real sum() const;
  Real absSum() const;
  // FIELD_3D.cpp:2960-2971, 3021-3032 on the GPU: NaN cells are skipped
  // unless cell 0 is NaN, ties keep the first cell
  Real fieldMax() const;
  Real fieldMin() const;
  // Of a signed distance field, on the GPU (FIELD_3D.cpp:5241-5297,
  // 5316-5352): 1 on the cells next to a sign change, then 1 over the
  // 6-neighbour layer number, negated inside and quartered outside; the
  // normalized version scales the inside and the outside to absolute sums
  // of 0.5 each. A field without a sign change throws (the reference sweeps
  // forever). absSum, when given, receives the result's absolute sum.
  FIELD_3D invertedDistance() const;
  FIELD_3D invertedDistanceNormalized(Real* absSum = nullptr) const;

 private:
  int _xRes = 0, _yRes = 0, _zRes = 0;
  VEC3F _center, _lengths{1, 1, 1};
  QUATERNION _rotation{1, 0, 0, 0};
  std::vector<Real> _data;
};

struct TRIANGLE {
  int v[3];
};

class TRIANGLE_MESH {
 public:
  TRIANGLE_MESH() {}
  // marching cubes on a field (TRIANGLE_MESH.cpp:48-57 -> computeStagedMarchingCubes)
  explicit TRIANGLE_MESH(const FIELD_3D& field, Real isosurface = 0.0);
  // Non-linear marching cubes on a field (TRIANGLE_MESH.cpp:62-79 ->
  // computeNonlinearMarchingCubes :152): cases from the field's signs,
  // vertices by bisection on the map re-evaluated on the GPU.
  TRIANGLE_MESH(const FIELD_3D& field, const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, Real expScaling,
                int maxIterations, Real slice, Real isosurface);
  // Non-linear low-memory mesher with an edge cache (TRIANGLE_MESH.cpp:84-101,
  // 681-768): slabs of the map evaluated on the fly, vertices refined by
  // bisection. When cacheFilename names an existing cache (readEdgeCacheHuge
  // layout) the mesh is rebuilt from it without evaluating slabs; otherwise
  // the cache is written after the run (an empty name skips the cache).
  // precision: slab evaluation (vertices are always FP64); chunkSlices:
  // z-slices evaluated per slab.
  TRIANGLE_MESH(const VEC3F& center, const VEC3F& lengths, const VEC3I& res, const POLYNOMIAL_4D& top,
                const POLYNOMIAL_4D& bottom, Real expScaling, int maxIterations, Real slice, Real isosurface,
                const QUATERNION& rotation, const std::string& cacheFilename, qj_precision precision = QJ_FP32_BAND,
                int chunkSlices = 64);
  // The same mesher with the Lagrange evaluation order (TRIANGLE_MESH.cpp:
  // 106-123, computeLagrangeNonlinearMarchingCubesLowMemoryHuge :773-857),
  // FP64 slabs.
  TRIANGLE_MESH(const VEC3F& center, const VEC3F& lengths, const VEC3I& res, const POLYNOMIAL_4D& top,
                const POLYNOMIAL_4D& bottom, Real expScaling, int maxIterations, Real slice, Real isosurface,
                const QUATERNION& rotation);

  std::vector<VEC3F>& vertices() { return _vertices; }
  const std::vector<VEC3F>& vertices() const { return _vertices; }
  std::vector<TRIANGLE>& triangles() { return _triangles; }
  const std::vector<TRIANGLE>& triangles() const { return _triangles; }
  bool writeOBJ(const std::string& filename) const;  // TRIANGLE_MESH.cpp:2762-2800

 private:
  std::vector<VEC3F> _vertices;
  std::vector<TRIANGLE> _triangles;
};

// The map / score / gradient members of OPTIMIZE_3D on the GPU.
class OPTIMIZE_3D {
 public:
  OPTIMIZE_3D();
  OPTIMIZE_3D(int xRes, int yRes, int zRes, const VEC3F& center, const VEC3F& lengths);
  ~OPTIMIZE_3D();
  OPTIMIZE_3D(const OPTIMIZE_3D&) = delete;
  OPTIMIZE_3D& operator=(const OPTIMIZE_3D&) = delete;

  POLYNOMIAL_4D& top() { return _top; }
  POLYNOMIAL_4D& bottom() { return _bottom; }
  FIELD_3D& fractal() { return _fractal; }
  FIELD_3D& weightField() { return _weightField; }
  FIELD_3D& distanceField() { return _distanceField; }
  FIELD_3D& inverseDistanceField() { return _inverseDistanceField; }
  Real inverseFieldSum() const { return _inverseFieldSum; }
  int& maxIterations() { return _maxIterations; }
  Real& expScaling() { return _expScaling; }
  Real& quaternionSlice() { return _quaternionSlice; }
  Real& tanhAlpha() { return _tanhAlpha; }
  Real& tanhThreshold() { return _tanhThreshold; }
  // precision of the field maps (QJ_FP64 or QJ_FP32_BAND) and the band width
  void setPrecision(qj_precision p, Real band = 0.05) { _precision = p; _band = band; }

  // OPTIMIZE_3D.cpp:1272-1370 (escape 2000, black-hole stop)
  void computeLogScaledPowerRationalMap(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, FIELD_3D& fractal,
                                        bool verbose = false);
  void computeLogScaledPowerRationalMap(bool verbose = false) {
    computeLogScaledPowerRationalMap(_top, _bottom, _fractal, verbose);
  }
  // OPTIMIZE_3D.cpp:1374-1568 (escape 20, origin skipped, overflowed cells re-passed)
  void computeRationalMapAP(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, FIELD_3D& fractal,
                            bool verbose = false);
  // OPTIMIZE_3D.cpp:1572-1718 (escape 20, cells at the origin keep their
  // value, no black-hole stop), FP64
  void computeLogScaledPowerRationalMapCached(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom,
                                              FIELD_3D& fractal, bool verbose = false);
  // computeLogPowerRationalMap(bool) -> computeLogPowerRationalMapCached
  // (OPTIMIZE_3D.h:164, OPTIMIZE_3D.cpp:1165-1269): the same map with the
  // power scalars left out (evaluatePowerFactored, POLYNOMIAL_4D.cpp:966-1003)
  // and no origin test, FP64
  void computeLogPowerRationalMap(bool verbose = false);

  // Scores (OPTIMIZE_3D.cpp:3573-3585, 4484-4563), sums on the GPU:
  // weightScore = sum over the cells of -tanh(alpha F) w(cell centre);
  // inverseDistanceScore the same with F - threshold (inverseSum unused, as
  // in the reference); the batched form scores many thresholds in one pass.
  Real computeLogScaledPowerRationalScore();
  Real computeMaximumScaledPowerRationalScore();
  Real weightScore(const FIELD_3D& logFractal, const FIELD_3D& weightField);
  Real inverseDistanceScore(const FIELD_3D& logFractal, const FIELD_3D& inverseDistanceField, const Real& inverseSum,
                            const Real& threshold);
  std::vector<Real> inverseDistanceScores(const FIELD_3D& logFractal, const FIELD_3D& inverseDistanceField,
                                          const std::vector<Real>& thresholds);
  // OPTIMIZE_3D.cpp:4694-4772: the unscaled one-iteration map at the
  // fractal's resolution, clampInfsToMean, then four scans of 100 thresholds
  // between the field's min and max for the lowest inverseDistanceScore
  // against invertedDistance(distance field); expScaling becomes
  // exp(-threshold). Returns the lowest score. A scan whose threshold stops
  // advancing throws (the reference would loop forever).
  Real estimateConformalRadius();

  // OPTIMIZE_3D.cpp:3442-3471: resolution, centre and lengths from the
  // distance field; the fractal becomes a zero field on that grid
  void setScoreFields(const FIELD_3D& distanceField, const FIELD_3D& curvatureField);
  void computeWeightField() { _weightField = _inverseDistanceField; }  // OPTIMIZE_3D.cpp:10840-10843
  // OPTIMIZE_3D.cpp:4338-4394: the problem scaled about the origin / moved
  void scaleEverything(const Real& scale);
  void translateEverything(const VEC3F& translation);
  void translateEverythingExceptFirst(const VEC3F& translation);

  // AP score and gradients at one iteration (OPTIMIZE_3D.cpp:3563, 3589, 8905, 10012, 9250, 10360).
  // One GPU pass computes all of them for the current state; the accessors
  // reuse it until the state changes.
  Real computeScoreAP();
  Real computeExpGradientAP();
  Real computeTopGradientAP(int whichRoot);
  Real computeBottomGradientAP(int whichRoot);
  Real computeTopBulkScalarGradientAP();
  Real computeBottomBulkScalarGradientAP();
  int adaptiveCells();

  // .o3d layout (OPTIMIZE_3D.cpp:3254-3346): int xRes, yRes, zRes; center;
  // lengths; top and bottom POLYNOMIAL_4D; double expScaling; int history
  // size (0, the history is never written); the fractal, distance and
  // curvature FIELD_3Ds. read() then does what the reference's read does
  // (:3348-3388): setScoreFields(distance, distance) (so the curvature field
  // becomes a copy of the distance field and the fractal a zero field on the
  // distance field's grid) and, unless an inverse distance field is already
  // set, inverseDistanceField = invertedDistanceNormalized(distance) with its
  // absolute sum, then weightField = inverseDistanceField.
  void write(const std::string& filename) const;
  void read(const std::string& filename);
  // the stored members only, without read()'s set-up (no GPU work): the
  // .o3d layout as written
  void readState(const std::string& filename);
  FIELD_3D& curvatureField() { return _curvatureField; }
  const VEC3F& center() const { return _center; }
  const VEC3F& lengths() const { return _lengths; }
  int xRes() const { return _xRes; }
  int yRes() const { return _yRes; }
  int zRes() const { return _zRes; }

 private:
  void sync_rational(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom);
  void scores_on_device(const FIELD_3D& logFractal, const FIELD_3D& weight, const std::vector<Real>& thresholds,
                        std::vector<Real>* scores, Real* maxScore);
  void map_into(const POLYNOMIAL_4D& top, const POLYNOMIAL_4D& bottom, FIELD_3D& fractal, const qj_map_params& p,
                bool keepSkipped);
  void run_ap();
  void ensure_device(size_t cells);

  int _xRes = 0, _yRes = 0, _zRes = 0;
  VEC3F _center, _lengths{1, 1, 1};
  POLYNOMIAL_4D _top, _bottom;
  FIELD_3D _fractal, _distanceField, _curvatureField, _weightField, _inverseDistanceField;
  Real _inverseFieldSum = 0;
  int _maxIterations = 1;
  Real _expScaling = 1.0;
  Real _quaternionSlice = 0.0;
  Real _tanhAlpha = 10.0, _tanhThreshold = 0.0;
  int _historySize = 0;
  qj_precision _precision = QJ_FP64;
  Real _band = 0.05;

  qj_rational* _rational = nullptr;
  qj_workspace* _ws = nullptr;
  double* _dField = nullptr;
  double* _dWeight = nullptr;
  size_t _dCells = 0, _dWeightCells = 0;
  bool _apValid = false;
  qj_ap_grads _ap{};
  std::vector<double> _gTop, _gBottom;
  std::vector<double> _apKey;
};

// TOUPEE callback ABI (TAO_SOLVER.h:41-43) with VECTOR = std::vector<Real>.
typedef std::vector<Real> VECTOR;
typedef void (*FunctionGradient)(const VECTOR& state, Real& value, VECTOR& gradient);
typedef void (*Monitor)();

// Bound-constrained limited-memory variable metric minimisation
// (TAO_SOLVER::optimizeBLMVM, TAO_SOLVER.cpp:72-341): projected quasi-Newton
// steps inside [lowerBound, upperBound] with a bounded More-Thuente line
// search, TAO's default convergence test and the reference's defaults
// (tolerances 1e-8, 2000 iterations).
class TAO_SOLVER {
 public:
  enum SOLVER_STATE {  // TAO_SOLVER.h:62-85
    CONVERGED_FATOL = 1,
    CONVERGED_FRTOL = 2,
    CONVERGED_GATOL = 3,
    CONVERGED_GRTOL = 4,
    CONVERGED_GTTOL = 5,
    CONVERGED_STEPTOL = 6,
    CONVERGED_MINF = 7,
    CONVERGED_USER = 8,
    DIVERGED_MAXITS = -2,
    DIVERGED_NAN = -4,
    DIVERGED_MAXFCN = -5,
    DIVERGED_LS_FAILURE = -6,
    CONTINUE_ITERATING = 0
  };

  TAO_SOLVER() {}
  void SetTolerances(Real fatol, Real frtol, Real gatol, Real grtol, Real gttol);
  int& maxIterations() { return _maxIterations; }
  Real& minimumValue() { return _fmin; }
  Real& maxScore() { return _maxScore; }  // progress printing scale in the reference; kept for callers

  // Minimises from `solution` (overwritten with the result); returns the
  // number of iterations. functionGradient is called once per trial point.
  int optimizeBLMVM(VECTOR& solution, const VECTOR& upperBound, const VECTOR& lowerBound,
                    FunctionGradient functionGradient, Monitor monitor = nullptr);

  SOLVER_STATE reason() const { return _reason; }
  int iterations() const { return _niter; }
  Real value() const { return _f; }
  // objective after the start and after each iteration
  const std::vector<Real>& history() const { return _history; }

 private:
  void converged(Real f, Real gnorm);

  Real _fatol = 1e-8, _frtol = 1e-8, _gatol = 1e-8, _grtol = 1e-8, _gttol = 0.0;
  Real _fmin = -1e100, _maxScore = 1.0;
  int _maxIterations = 2000;
  int _niter = 0;
  Real _gnorm0 = 0, _f = 0;
  SOLVER_STATE _reason = CONTINUE_ITERATING;
  std::vector<Real> _history;
};

// The shape-optimizer driver of projects/toothOptimizer/toothOptimizer.cpp (a
// program in the reference, a function here): read() has already set the
// weight field; the set-up of its main() (:498-520) moves and scales the
// problem to the fractal radius, multiplies the root powers by the initial
// amplitude and resets both power scalars to 1; then optimizeSpecificResolution
// (:385-457) runs for each resolution: at 25 the conformal radius estimate and
// the fractal mesh, the initial and maximum scores, a BLMVM solve over
// (log c, top scalar, bottom scalar) then one over every exponent
// (FormScaledOnlyFunctionGradient / FormScaledFunctionGradient, :64-134, with
// the GPU AP pass as the callback), a solve with the inside-reweighted weight
// field (buildReweightedInsideField, :276-381), a last solve with the
// original weights, and the state written to <path><outputFilename>_res_<res>.o3d.
// monitor (:48-61) writes <path>output.o3d on every checkpointEvery-th BLMVM
// iteration. The reweighting needs the signed distance field of the fractal's
// mesh, which the reference gets from an external program (sdfGen on
// temp.obj); here the caller supplies it (fractalDistance), and without it
// the reweighted solve is skipped. One schedule runs at a time per process
// (the callbacks reach the optimizer through process-wide state, as the
// reference's globals do).
struct ToothOptimizerSettings {
  std::string path = "./temp/";                 // "path"
  std::string outputFilename = "temp.o3d";      // "optimization output filename"
  std::vector<int> resolutions{25, 50, 97};
  Real fractalRadius = 4.0;                     // "fractal radius"
  Real initialBulkScalarAmplitude = -1;         // "initial bulk scalar amplitude"
  int maxTaoIterations = 2000;                  // TAO's default (the reference parses but never applies its own)
  int checkpointEvery = 10;
  bool prepare = true;                          // run main()'s set-up first
  // the fractal mesh's signed distance on the grid of meshDistance (the
  // optimizer's distance field), one value per cell
  std::function<FIELD_3D(const TRIANGLE_MESH& fractalMesh, const FIELD_3D& meshDistance)> fractalDistance;
};

struct ToothSolveRecord {
  std::string name;  // "scaled-only", "scaled", "reweighted", "unweighted"
  int iterations = 0, reason = 0, evaluations = 0;
  Real value = 0, maxScore = 0;
  std::vector<Real> solution, history;
};

struct ToothResolutionRecord {
  int res = 0;
  bool conformal = false;
  Real conformalScore = 0, conformalExpScaling = 0;
  int64_t fractalVertices = 0, fractalTriangles = 0;      // computeFractal at resolution 25
  int64_t reweightVertices = 0, reweightTriangles = 0;    // the mesh the reweighting starts from
  Real initialScore = 0, maxScore = 0;
  std::vector<ToothSolveRecord> solves;
  std::string output;
};

struct ToothOptimizerLog {
  std::vector<ToothResolutionRecord> resolutions;
  int checkpoints = 0;
};

void optimizeTooth(OPTIMIZE_3D& optimize3D, const ToothOptimizerSettings& settings, ToothOptimizerLog* log = nullptr);

}  // namespace quijibo
