// Test program for the C++ host layer (include/quijibo/quijibo.hpp), run by
// tests/test_formats.py (io and copy modes, no GPU) and
// tests/test_gpu_cpp_host_layer.py (map / ap modes, GPU).
//
//   host_layer_driver io  <dir>
//       writes a POLYNOMIAL_4D, FIELD_3D (plain and .gz) and .o3d file into
//       <dir>, reads each back and checks the round trip field by field.
//   host_layer_driver copy <in.o3d> <out.o3d>
//       OPTIMIZE_3D::readState then OPTIMIZE_3D::write (the layout round trip).
//   host_layer_driver map <in.o3d> <maxIterations> <precision 0|1> <outdir>
//       OPTIMIZE_3D::read, computeLogScaledPowerRationalMap into the fractal
//       -> <outdir>/field.bin; TRIANGLE_MESH(fractal) -> linear.obj; the
//       fractal with clampInfsToMean and setBorders(totalCells) -> clamped.bin
//       and its mesh -> clamped.obj;
//       TRIANGLE_MESH(center, lengths, res, top, bottom, ..., cache) -> nonlinear.obj
//       (writes <outdir>/mesh.cache), again from the cache -> resumed.obj, and
//       the Lagrange mesher -> lagrange.obj.
//   host_layer_driver ap <in.o3d> <weight.field> <alpha> <threshold>
//       one-iteration AP score and gradients printed as JSON.
//   host_layer_driver blmvm rosenbrock|boxquad
//       TAO_SOLVER::optimizeBLMVM on a test problem (no GPU); JSON result.
//   host_layer_driver readsetup <in.o3d> <outdir>
//       OPTIMIZE_3D::read with its set-up; weight, inverse distance,
//       curvature and fractal fields -> <outdir>; JSON summary.
//   host_layer_driver tooth <in.o3d> <path/> <maxTao> <amplitude> <radius> <res,res,...>
//       the toothOptimizer schedule (optimizeTooth) with a sphere of radius 1
//       as the fractal's distance field; JSON log.
//   host_layer_driver optimize <in.o3d> <weight.field> <alpha> <iterations>
//       the shape optimizer's exponent fit (toothOptimizer.cpp:65-110,
//       160-218): BLMVM over [log c, bottom powers, top powers, top scalar,
//       bottom scalar] with the GPU AP score and gradients as the callback.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>

#include "quijibo/quijibo.hpp"

using namespace quijibo;

static int fail(const char* what) {
  fprintf(stderr, "FAIL: %s\n", what);
  return 1;
}

static bool same_field(const FIELD_3D& a, const FIELD_3D& b) {
  if (a.xRes() != b.xRes() || a.yRes() != b.yRes() || a.zRes() != b.zRes()) return false;
  for (int i = 0; i < 3; i++)
    if (a.center()[i] != b.center()[i] || a.lengths()[i] != b.lengths()[i]) return false;
  return memcmp(a.dataConst(), b.dataConst(), sizeof(double) * a.totalCells()) == 0;
}

static bool same_poly(const POLYNOMIAL_4D& a, const POLYNOMIAL_4D& b) {
  if (a.totalRoots() != b.totalRoots() || a.powerScalar() != b.powerScalar()) return false;
  for (int i = 0; i < a.totalRoots(); i++) {
    if (a.powers()[i] != b.powers()[i]) return false;
    for (int k = 0; k < 4; k++)
      if (a.roots()[i][k] != b.roots()[i][k]) return false;
  }
  return true;
}

static int run_io(const std::string& dir) {
  POLYNOMIAL_4D top;
  top.addRoot(QUATERNION(0.5, -0.25, 0.125, 0.0), 1.0);
  top.addRoot(QUATERNION(-0.75, 0.5, 0.3, 0.01), 2.0);
  top.powerScalar() = 1.25;
  POLYNOMIAL_4D bottom(std::vector<QUATERNION>{QUATERNION(0.1, 0.2, 0.3, 0.4)});
  FILE* f = fopen((dir + "/poly.bin").c_str(), "wb");
  if (!f) return fail("open poly.bin");
  top.write(f);
  fclose(f);
  f = fopen((dir + "/poly.bin").c_str(), "rb");
  POLYNOMIAL_4D back;
  back.read(f);
  fclose(f);
  if (!same_poly(top, back)) return fail("POLYNOMIAL_4D round trip");

  FIELD_3D field(5, 4, 3, VEC3F(0.1, -0.2, 0.3), VEC3F(3.0, 2.0, 1.5));
  for (size_t i = 0; i < field.totalCells(); i++) field[i] = 0.5 * (double)i - 7.0 + 1e-3 * (double)(i * i);
  field.write(dir + "/field.bin");
  field.write(dir + "/field.bin.gz");
  FIELD_3D a, b;
  a.read(dir + "/field.bin");
  b.read(dir + "/field.bin.gz");
  if (!same_field(field, a)) return fail("FIELD_3D round trip");
  if (!same_field(field, b)) return fail("FIELD_3D gz round trip");

  OPTIMIZE_3D opt(5, 4, 3, VEC3F(0.1, -0.2, 0.3), VEC3F(3.0, 2.0, 1.5));
  opt.top() = top;
  opt.bottom() = bottom;
  opt.expScaling() = 0.75;
  opt.fractal() = field;
  opt.distanceField() = field;
  opt.distanceField()[0] = 42.0;
  opt.write(dir + "/state.o3d");
  OPTIMIZE_3D in;
  in.readState(dir + "/state.o3d");
  if (in.xRes() != 5 || in.yRes() != 4 || in.zRes() != 3) return fail(".o3d resolution");
  if (!same_poly(in.top(), top) || !same_poly(in.bottom(), bottom)) return fail(".o3d polynomials");
  if (in.expScaling() != 0.75) return fail(".o3d expScaling");
  if (!same_field(in.fractal(), field) || !same_field(in.distanceField(), opt.distanceField()))
    return fail(".o3d fields");
  if (in.curvatureField().totalCells() != 0) return fail(".o3d curvature field");

  // an invalid file is an error, not an exit
  FILE* bad = fopen((dir + "/short.o3d").c_str(), "wb");
  int three[3] = {4, 4, 4};
  fwrite(three, sizeof(int), 3, bad);
  fclose(bad);
  bool threw = false;
  try {
    OPTIMIZE_3D s;
    s.readState(dir + "/short.o3d");
  } catch (const Error& e) {
    threw = e.status == QJ_ERR_INVALID;
  }
  if (!threw) return fail("short .o3d not rejected");
  printf("io ok\n");
  return 0;
}

static int run_map(const std::string& in, int maxIter, int precision, const std::string& out) {
  OPTIMIZE_3D opt;
  opt.read(in);
  opt.maxIterations() = maxIter;
  opt.setPrecision(precision ? QJ_FP32_BAND : QJ_FP64);
  FIELD_3D& fractal = opt.fractal();
  opt.computeLogScaledPowerRationalMap();
  fractal.write(out + "/field.bin");
  TRIANGLE_MESH linear(fractal);
  if (!linear.writeOBJ(out + "/linear.obj")) return fail("writeOBJ linear");
  // the reference's clean-up before meshing a fractal (toothOptimizer.cpp:250-254)
  FIELD_3D fieldCopy(fractal);
  fieldCopy.clampInfsToMean();
  fieldCopy.setBorders((Real)fieldCopy.totalCells());
  fieldCopy.write(out + "/clamped.bin");
  TRIANGLE_MESH clamped(fieldCopy);
  if (!clamped.writeOBJ(out + "/clamped.obj")) return fail("writeOBJ clamped");
  const VEC3I res(opt.xRes(), opt.yRes(), opt.zRes());
  const std::string cache = out + "/mesh.cache";
  // first run writes the edge cache, the second rebuilds the mesh from it
  TRIANGLE_MESH nonlinear(opt.center(), opt.lengths(), res, opt.top(), opt.bottom(), opt.expScaling(), maxIter,
                          opt.quaternionSlice(), 0.0, QUATERNION(1, 0, 0, 0), cache,
                          precision ? QJ_FP32_BAND : QJ_FP64, 7);
  if (!nonlinear.writeOBJ(out + "/nonlinear.obj")) return fail("writeOBJ nonlinear");
  TRIANGLE_MESH resumed(opt.center(), opt.lengths(), res, opt.top(), opt.bottom(), opt.expScaling(), maxIter,
                        opt.quaternionSlice(), 0.0, QUATERNION(1, 0, 0, 0), cache);
  if (!resumed.writeOBJ(out + "/resumed.obj")) return fail("writeOBJ resumed");
  TRIANGLE_MESH lagrange(opt.center(), opt.lengths(), res, opt.top(), opt.bottom(), opt.expScaling(), maxIter,
                         opt.quaternionSlice(), 0.0, QUATERNION(1, 0, 0, 0));
  if (!lagrange.writeOBJ(out + "/lagrange.obj")) return fail("writeOBJ lagrange");
  printf("map ok %zu %zu %zu %zu %zu\n", linear.vertices().size(), linear.triangles().size(),
         nonlinear.vertices().size(), nonlinear.triangles().size(), lagrange.triangles().size());
  return 0;
}

static int run_ap(const std::string& in, const std::string& weight, double alpha, double thr) {
  OPTIMIZE_3D opt;
  opt.read(in);
  opt.maxIterations() = 1;
  opt.tanhAlpha() = alpha;
  opt.tanhThreshold() = thr;
  opt.weightField().read(weight);
  const double score = opt.computeScoreAP();
  printf("{\"score\": %.17g, \"grad_exp\": %.17g, \"grad_top_scalar\": %.17g, \"grad_bottom_scalar\": %.17g, ",
         score, opt.computeExpGradientAP(), opt.computeTopBulkScalarGradientAP(),
         opt.computeBottomBulkScalarGradientAP());
  printf("\"adaptive_cells\": %d, \"grad_top\": [", opt.adaptiveCells());
  for (int i = 0; i < opt.top().totalRoots(); i++) printf("%s%.17g", i ? ", " : "", opt.computeTopGradientAP(i));
  printf("], \"grad_bottom\": [");
  for (int i = 0; i < opt.bottom().totalRoots(); i++)
    printf("%s%.17g", i ? ", " : "", opt.computeBottomGradientAP(i));
  printf("]}\n");
  return 0;
}

// ---- BLMVM test problems (the same functions as tests/test_blmvm.py) ----
static void rosenbrock(const VECTOR& x, Real& f, VECTOR& g) {
  const size_t n = x.size();
  f = 0;
  g.assign(n, 0.0);
  for (size_t i = 0; i + 1 < n; i++) {
    const Real a = x[i + 1] - x[i] * x[i], b = 1.0 - x[i];
    f += 100.0 * a * a + b * b;
    g[i] += -400.0 * a * x[i] - 2.0 * b;
    g[i + 1] += 200.0 * a;
  }
}

static void boxquad(const VECTOR& x, Real& f, VECTOR& g) {
  // 0.5 x'Ax - b'x with A = tridiag(-1, 2.5, -1)
  static const Real b[6] = {1.0, -2.0, 3.0, 0.5, -1.0, 2.0};
  const size_t n = x.size();
  f = 0;
  g.assign(n, 0.0);
  for (size_t i = 0; i < n; i++) {
    Real ax = 2.5 * x[i];
    if (i > 0) ax -= x[i - 1];
    if (i + 1 < n) ax -= x[i + 1];
    g[i] = ax - b[i];
    f += 0.5 * x[i] * ax - b[i] * x[i];
  }
}

static void print_vec(const char* name, const VECTOR& v) {
  printf("\"%s\": [", name);
  for (size_t i = 0; i < v.size(); i++) printf("%s%.17g", i ? ", " : "", v[i]);
  printf("]");
}

static int run_blmvm(const std::string& problem) {
  TAO_SOLVER tao;
  VECTOR x, lo, hi;
  FunctionGradient fg;
  if (problem == "rosenbrock") {
    x = {-1.2, 1.0, -1.2, 1.0};
    lo.assign(4, -3.4028234663852886e38);
    hi.assign(4, 3.4028234663852886e38);
    fg = rosenbrock;
  } else if (problem == "boxquad") {
    x.assign(6, 0.0);
    lo.assign(6, -0.5);
    hi.assign(6, 0.8);
    fg = boxquad;
  } else {
    return fail("unknown problem");
  }
  tao.optimizeBLMVM(x, hi, lo, fg);
  printf("{\"iterations\": %d, \"reason\": %d, \"f\": %.17g, ", tao.iterations(), (int)tao.reason(), tao.value());
  print_vec("x", x);
  printf(", ");
  print_vec("history", tao.history());
  printf("}\n");
  return 0;
}

// ---- the exponent fit with the GPU callback ----
static OPTIMIZE_3D* g_opt = nullptr;
static int g_evals = 0;

static void scaled_function_gradient(const VECTOR& state, Real& function, VECTOR& gradient) {
  OPTIMIZE_3D& o = *g_opt;
  const int nt = o.top().totalRoots(), nb = o.bottom().totalRoots();
  const int scalingStart = 1 + nt + nb;
  o.expScaling() = exp(state[0]);
  int k = 1;
  for (int i = 0; i < nb; i++, k++) o.bottom().powersMutable()[i] = state[k];
  for (int i = 0; i < nt; i++, k++) o.top().powersMutable()[i] = state[k];
  o.top().powerScalar() = state[scalingStart];
  o.bottom().powerScalar() = state[scalingStart + 1];
  function = o.computeScoreAP();
  gradient.resize(scalingStart + 2);
  gradient[0] = o.computeExpGradientAP();
  for (int i = 0; i < nb; i++) gradient[i + 1] = o.computeBottomGradientAP(i);
  for (int i = 0; i < nt; i++) gradient[i + 1 + nb] = o.computeTopGradientAP(i);
  gradient[scalingStart] = o.computeTopBulkScalarGradientAP();
  gradient[scalingStart + 1] = o.computeBottomBulkScalarGradientAP();
  g_evals++;
}

static int run_optimize(const std::string& in, const std::string& weight, double alpha, int iterations) {
  OPTIMIZE_3D opt;
  opt.read(in);
  opt.maxIterations() = 1;
  opt.tanhAlpha() = alpha;
  opt.weightField().read(weight);
  g_opt = &opt;
  const int nt = opt.top().totalRoots(), nb = opt.bottom().totalRoots();
  const int n = 1 + nt + nb + 2;
  VECTOR x(n);
  x[0] = log(opt.expScaling());
  int k = 1;
  for (int i = 0; i < nb; i++, k++) x[k] = opt.bottom().powers()[i];
  for (int i = 0; i < nt; i++, k++) x[k] = opt.top().powers()[i];
  x[n - 2] = opt.top().powerScalar();
  x[n - 1] = opt.bottom().powerScalar();
  VECTOR hi(n, 3.4028234663852886e38), lo(n, 0.0);  // toothOptimizer.cpp:189-191
  lo[0] = -3.4028234663852886e38;
  TAO_SOLVER tao;
  tao.maxIterations() = iterations;
  tao.optimizeBLMVM(x, hi, lo, scaled_function_gradient);
  printf("{\"iterations\": %d, \"reason\": %d, \"evaluations\": %d, \"f\": %.17g, ", tao.iterations(),
         (int)tao.reason(), g_evals, tao.value());
  print_vec("x", x);
  printf(", ");
  print_vec("history", tao.history());
  printf("}\n");
  return 0;
}

// read() with the reference's set-up: the weight and inverse distance fields
// and the fractal's new grid go to <outdir>
static int run_readsetup(const std::string& in, const std::string& out) {
  OPTIMIZE_3D opt;
  opt.read(in);
  opt.weightField().write(out + "/weight.bin");
  opt.inverseDistanceField().write(out + "/inverse.bin");
  opt.curvatureField().write(out + "/curvature.bin");
  opt.fractal().write(out + "/fractal.bin");
  printf("{\"res\": [%d, %d, %d], \"inverse_sum\": %.17g}\n", opt.xRes(), opt.yRes(), opt.zRes(),
         opt.inverseFieldSum());
  return 0;
}

// the fractal distance the reweighted solve uses: a sphere of radius 1 about
// the origin, on the grid of the optimizer's distance field
static FIELD_3D sphere_distance(const TRIANGLE_MESH& mesh, const FIELD_3D& grid) {
  (void)mesh;
  FIELD_3D f(grid);
  for (int z = 0; z < f.zRes(); z++)
    for (int y = 0; y < f.yRes(); y++)
      for (int x = 0; x < f.xRes(); x++) {
        const VEC3F c = f.cellCenter(x, y, z);
        f(x, y, z) = sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) - 1.0;
      }
  return f;
}

static int run_tooth(const std::string& in, const std::string& path, int maxTao, double amp, double radius,
                     const std::string& resList) {
  OPTIMIZE_3D opt;
  opt.read(in);
  ToothOptimizerSettings s;
  s.path = path;
  s.outputFilename = "tooth.o3d";
  s.resolutions.clear();
  for (size_t a = 0; a < resList.size();) {
    size_t b = resList.find(',', a);
    if (b == std::string::npos) b = resList.size();
    s.resolutions.push_back(atoi(resList.substr(a, b - a).c_str()));
    a = b + 1;
  }
  s.maxTaoIterations = maxTao;
  s.initialBulkScalarAmplitude = amp;
  s.fractalRadius = radius;
  s.fractalDistance = sphere_distance;
  ToothOptimizerLog log;
  optimizeTooth(opt, s, &log);
  printf("{\"checkpoints\": %d, \"resolutions\": [", log.checkpoints);
  for (size_t i = 0; i < log.resolutions.size(); i++) {
    const ToothResolutionRecord& r = log.resolutions[i];
    printf("%s{\"res\": %d, \"conformal\": %s, \"conformal_score\": %.17g, \"conformal_exp_scaling\": %.17g, "
           "\"fractal_vertices\": %lld, \"fractal_triangles\": %lld, \"reweight_triangles\": %lld, "
           "\"initial_score\": %.17g, "
           "\"max_score\": %.17g, \"output\": \"%s\", \"solves\": [",
           i ? ", " : "", r.res, r.conformal ? "true" : "false", r.conformalScore, r.conformalExpScaling,
           (long long)r.fractalVertices, (long long)r.fractalTriangles, (long long)r.reweightTriangles, r.initialScore,
           r.maxScore, r.output.c_str());
    for (size_t k = 0; k < r.solves.size(); k++) {
      const ToothSolveRecord& v = r.solves[k];
      printf("%s{\"name\": \"%s\", \"iterations\": %d, \"reason\": %d, \"evaluations\": %d, \"value\": %.17g, "
             "\"max_score\": %.17g, ",
             k ? ", " : "", v.name.c_str(), v.iterations, v.reason, v.evaluations, v.value, v.maxScore);
      print_vec("x", v.solution);
      printf(", ");
      print_vec("history", v.history);
      printf("}");
    }
    printf("]}");
  }
  printf("]}\n");
  return 0;
}

int main(int argc, char** argv) {
  try {
    const std::string mode = argc > 1 ? argv[1] : "";
    if (mode == "io" && argc == 3) return run_io(argv[2]);
    if (mode == "copy" && argc == 4) {
      OPTIMIZE_3D o;
      o.readState(argv[2]);
      o.write(argv[3]);
      printf("copy ok\n");
      return 0;
    }
    if (mode == "map" && argc == 6) return run_map(argv[2], atoi(argv[3]), atoi(argv[4]), argv[5]);
    if (mode == "ap" && argc == 6) return run_ap(argv[2], argv[3], atof(argv[4]), atof(argv[5]));
    if (mode == "blmvm" && argc == 3) return run_blmvm(argv[2]);
    if (mode == "optimize" && argc == 6) return run_optimize(argv[2], argv[3], atof(argv[4]), atoi(argv[5]));
    if (mode == "readsetup" && argc == 4) return run_readsetup(argv[2], argv[3]);
    if (mode == "tooth" && argc == 8)
      return run_tooth(argv[2], argv[3], atoi(argv[4]), atof(argv[5]), atof(argv[6]), argv[7]);
  } catch (const Error& e) {
    fprintf(stderr, "quijibo::Error %d: %s\n", e.status, e.what());
    return 2;
  }
  fprintf(stderr, "usage: host_layer_driver io <dir> | map <o3d> <maxIter> <prec> <outdir> | ap <o3d> <weight> "
                  "<alpha> <thr>\n");
  return 64;
}
