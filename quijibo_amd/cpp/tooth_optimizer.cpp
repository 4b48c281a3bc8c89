// The shape-optimizer schedule of projects/toothOptimizer/toothOptimizer.cpp
// over the GPU OPTIMIZE_3D: the objective and every gradient of each BLMVM
// evaluation come from one fused AP pass on the device (qj_ap_score_gradients),
// the conformal-radius scan scores each pass's thresholds in one launch
// (qj_weight_scores), and the weight fields come from the device inverted
// distance. What stays on the host is the reference's bookkeeping: packing
// the state vector, the once-per-resolution reweighting arithmetic, and files.
#include <float.h>
#include <math.h>
#include <stdio.h>

#include "../../include/quijibo/quijibo.hpp"

namespace quijibo {

namespace {

// toothOptimizer.cpp keeps the optimizer, the path and the monitor counter as
// globals; the BLMVM callbacks are plain function pointers, so they do here too.
struct Schedule {
  OPTIMIZE_3D* opt = nullptr;
  const ToothOptimizerSettings* settings = nullptr;
  int monitorCounter = 0;
  int evaluations = 0;
  int checkpoints = 0;
};
Schedule g_run;

// monitor (toothOptimizer.cpp:48-61)
void monitor() {
  const int every = g_run.settings->checkpointEvery > 0 ? g_run.settings->checkpointEvery : 10;
  if (g_run.monitorCounter % every == 0) {
    g_run.opt->write(g_run.settings->path + "output.o3d");
    g_run.checkpoints++;
  }
  g_run.monitorCounter++;
}

// FormScaledFunctionGradient (toothOptimizer.cpp:64-110): state = (log c,
// bottom powers, top powers, top scalar, bottom scalar)
void formScaledFunctionGradient(const VECTOR& state, Real& function, VECTOR& gradient) {
  OPTIMIZE_3D& o = *g_run.opt;
  const int nt = o.top().totalRoots(), nb = o.bottom().totalRoots();
  const int scalingStart = 1 + nt + nb;
  o.expScaling() = exp(state[0]);
  int k = 1;
  for (int i = 0; i < nb; i++, k++) o.bottom().powersMutable()[i] = state[k];
  for (int i = 0; i < nt; i++, k++) o.top().powersMutable()[i] = state[k];
  o.top().powerScalar() = state[scalingStart];
  o.bottom().powerScalar() = state[scalingStart + 1];
  function = o.computeScoreAP();
  g_run.evaluations++;
  gradient.assign(scalingStart + 2, 0.0);
  gradient[0] = o.computeExpGradientAP();
  for (int i = 0; i < nb; i++) gradient[i + 1] = o.computeBottomGradientAP(i);
  for (int i = 0; i < nt; i++) gradient[i + 1 + nb] = o.computeTopGradientAP(i);
  gradient[scalingStart] = o.computeTopBulkScalarGradientAP();
  gradient[scalingStart + 1] = o.computeBottomBulkScalarGradientAP();
}

// FormScaledOnlyFunctionGradient (toothOptimizer.cpp:114-134): state = (log c,
// top scalar, bottom scalar)
void formScaledOnlyFunctionGradient(const VECTOR& state, Real& function, VECTOR& gradient) {
  OPTIMIZE_3D& o = *g_run.opt;
  o.expScaling() = exp(state[0]);
  o.top().powerScalar() = state[1];
  o.bottom().powerScalar() = state[2];
  function = o.computeScoreAP();
  g_run.evaluations++;
  gradient.assign(3, 0.0);
  gradient[0] = o.computeExpGradientAP();
  gradient[1] = o.computeTopBulkScalarGradientAP();
  gradient[2] = o.computeBottomBulkScalarGradientAP();
}

ToothSolveRecord runTao(const std::string& name, VECTOR& solution, FunctionGradient fg, Real maxScore) {
  // bounds (toothOptimizer.cpp:150-152, 190-192): [0, FLT_MAX], log c free
  VECTOR upper(solution.size(), (Real)FLT_MAX), lower(solution.size(), 0.0);
  lower[0] = -(Real)FLT_MAX;
  TAO_SOLVER tao;
  tao.SetTolerances(1e-8, 1e-8, 1e-8, 1e-8, 0);
  tao.maxScore() = maxScore;
  tao.maxIterations() = g_run.settings->maxTaoIterations;
  const int before = g_run.evaluations;
  tao.optimizeBLMVM(solution, upper, lower, fg, monitor);
  ToothSolveRecord r;
  r.name = name;
  r.iterations = tao.iterations();
  r.reason = (int)tao.reason();
  r.evaluations = g_run.evaluations - before;
  r.value = tao.value();
  r.maxScore = maxScore;
  r.solution = solution;
  r.history = tao.history();
  return r;
}

// callConstrainedScaledOnlyTao (toothOptimizer.cpp:138-169)
ToothSolveRecord scaledOnlySolve(OPTIMIZE_3D& o, Real maxScore) {
  VECTOR x{log(o.expScaling()), o.top().powerScalar(), o.bottom().powerScalar()};
  ToothSolveRecord r = runTao("scaled-only", x, formScaledOnlyFunctionGradient, maxScore);
  o.expScaling() = exp(x[0]);
  o.top().powerScalar() = x[1];
  o.bottom().powerScalar() = x[2];
  return r;
}

// callConstrainedScaledTao (toothOptimizer.cpp:173-220)
ToothSolveRecord scaledSolve(OPTIMIZE_3D& o, const std::string& name, Real maxScore) {
  const int nt = o.top().totalRoots(), nb = o.bottom().totalRoots();
  const int scalingStart = 1 + nt + nb;
  VECTOR x(scalingStart + 2);
  x[0] = log(o.expScaling());
  int k = 1;
  for (int i = 0; i < nb; i++, k++) x[k] = o.bottom().powers()[i];
  for (int i = 0; i < nt; i++, k++) x[k] = o.top().powers()[i];
  x[scalingStart] = o.top().powerScalar();
  x[scalingStart + 1] = o.bottom().powerScalar();
  ToothSolveRecord r = runTao(name, x, formScaledFunctionGradient, maxScore);
  o.expScaling() = exp(x[0]);
  k = 1;
  for (int i = 0; i < nb; i++, k++) o.bottom().powersMutable()[i] = x[k];
  for (int i = 0; i < nt; i++, k++) o.top().powersMutable()[i] = x[k];
  o.top().powerScalar() = x[scalingStart];
  o.bottom().powerScalar() = x[scalingStart + 1];
  return r;
}

// computeFractal (toothOptimizer.cpp:240-272): the one-iteration map on the
// fractal's grid, cleaned up and meshed (the res argument is only printed there)
TRIANGLE_MESH computeFractal(OPTIMIZE_3D& o) {
  o.maxIterations() = 1;
  o.computeLogScaledPowerRationalMap(true);
  FIELD_3D fieldCopy(o.fractal());
  fieldCopy.clampInfsToMean();
  fieldCopy.setBorders((Real)fieldCopy.totalCells());
  return TRIANGLE_MESH(fieldCopy);
}

// buildReweightedInsideField (toothOptimizer.cpp:276-381), with the fractal's
// signed distance from the caller instead of sdfGen
void buildReweightedInsideField(OPTIMIZE_3D& o, const ToothOptimizerSettings& s, ToothResolutionRecord& rec) {
  const TRIANGLE_MESH mesh = computeFractal(o);
  rec.reweightVertices = (int64_t)mesh.vertices().size();
  rec.reweightTriangles = (int64_t)mesh.triangles().size();
  const FIELD_3D& meshDistance = o.distanceField();
  FIELD_3D fractalDistance = s.fractalDistance(mesh, meshDistance);
  if (fractalDistance.totalCells() != meshDistance.totalCells() || o.weightField().totalCells() != meshDistance.totalCells())
    throw Error(QJ_ERR_INVALID, "buildReweightedInsideField: the fractal distance, the distance field and the "
                                "weight field must have the same number of cells");
  // the fractal distance's own grid, as the reference places it (only its mesh uses it)
  for (int i = 0; i < 3; i++) fractalDistance.center()[i] += -0.5;
  fractalDistance.scaleLengths(s.fractalRadius);
  for (int i = 0; i < 3; i++) fractalDistance.center()[i] *= s.fractalRadius;

  const size_t n = meshDistance.totalCells();
  FIELD_3D reweighted(o.weightField());
  Real wellSum = 0, poorSum = 0;
  for (size_t x = 0; x < n; x++)
    if (meshDistance[x] < 0) {
      const bool poor = fractalDistance[x] > 0, well = fractalDistance[x] < 0;
      if (poor) poorSum += reweighted[x];
      if (well) wellSum += reweighted[x];
    }
  wellSum = fabs(wellSum);
  poorSum = fabs(poorSum);
  int totalOutside = 0;
  for (size_t x = 0; x < n; x++)
    if (meshDistance[x] > 0) totalOutside++;
  const Real totalWorth = 0.5, outsideWorth = 0.5;
  const Real outsidePerCell = ((outsideWorth * totalWorth) / (1.0 - outsideWorth)) / totalOutside;
  const Real poorFraction = 0.5, wellFraction = 1.0 - poorFraction;
  const Real poorReweight = totalWorth * poorFraction / poorSum;
  const Real wellReweight = totalWorth * wellFraction / wellSum;
  for (size_t x = 0; x < n; x++) {
    if (meshDistance[x] < 0) {
      if (fractalDistance[x] > 0 && meshDistance[x] < 0) reweighted[x] *= poorReweight;
      if (fractalDistance[x] < 0 && meshDistance[x] < 0) reweighted[x] *= wellReweight;
    } else {
      reweighted[x] = outsidePerCell;
    }
  }
  o.weightField() = reweighted;
}

// optimizeSpecificResolution (toothOptimizer.cpp:385-457)
void optimizeSpecificResolution(OPTIMIZE_3D& o, const ToothOptimizerSettings& s, int res, ToothResolutionRecord& rec) {
  rec.res = res;
  const VEC3F center = o.fractal().center(), lengths = o.fractal().lengths();
  o.fractal() = FIELD_3D(res, res, res, center, lengths);
  o.maxIterations() = 1;
  if (res == 25) {
    rec.conformal = true;
    rec.conformalScore = o.estimateConformalRadius();
    rec.conformalExpScaling = o.expScaling();
    const TRIANGLE_MESH mesh = computeFractal(o);
    rec.fractalVertices = (int64_t)mesh.vertices().size();
    rec.fractalTriangles = (int64_t)mesh.triangles().size();
  }
  rec.initialScore = o.computeLogScaledPowerRationalScore();
  Real maxScore = o.computeMaximumScaledPowerRationalScore();
  rec.maxScore = maxScore;
  rec.solves.push_back(scaledOnlySolve(o, maxScore));
  rec.solves.push_back(scaledSolve(o, "scaled", maxScore));
  const FIELD_3D originalWeights = o.weightField();
  if (s.fractalDistance) {
    buildReweightedInsideField(o, s, rec);
    maxScore = o.computeMaximumScaledPowerRationalScore();
    rec.solves.push_back(scaledSolve(o, "reweighted", maxScore));
  }
  o.weightField() = originalWeights;
  maxScore = o.computeMaximumScaledPowerRationalScore();
  rec.solves.push_back(scaledSolve(o, "unweighted", maxScore));
  char suffix[64];
  snprintf(suffix, sizeof suffix, "_res_%i.o3d", res);
  rec.output = s.path + s.outputFilename + suffix;
  o.write(rec.output);
}

}  // namespace

void optimizeTooth(OPTIMIZE_3D& optimize3D, const ToothOptimizerSettings& settings, ToothOptimizerLog* log) {
  if (settings.fractalRadius <= 0) throw Error(QJ_ERR_INVALID, "optimizeTooth: fractal radius must be > 0");
  g_run = Schedule();
  g_run.opt = &optimize3D;
  g_run.settings = &settings;
  ToothOptimizerLog local;
  ToothOptimizerLog& out = log ? *log : local;
  out = ToothOptimizerLog();
  try {
    if (settings.prepare) {
      // toothOptimizer.cpp:498-520
      if (optimize3D.lengths()[0] < 0.95 * settings.fractalRadius) {
        optimize3D.translateEverythingExceptFirst(VEC3F(-0.5, -0.5, -0.5));
        optimize3D.scaleEverything(settings.fractalRadius);
      }
      const Real amp = settings.initialBulkScalarAmplitude;
      POLYNOMIAL_4D& top = optimize3D.top();
      POLYNOMIAL_4D& bottom = optimize3D.bottom();
      for (int x = 1; x < top.totalRoots(); x++) top.powersMutable()[x] *= amp;
      for (int x = 1; x < bottom.totalRoots(); x++) bottom.powersMutable()[x] *= amp;
      top.powerScalar() = 1;
      bottom.powerScalar() = 1;
    }
    for (int res : settings.resolutions) {
      out.resolutions.emplace_back();
      optimizeSpecificResolution(optimize3D, settings, res, out.resolutions.back());
    }
  } catch (...) {
    out.checkpoints = g_run.checkpoints;
    g_run = Schedule();
    throw;
  }
  out.checkpoints = g_run.checkpoints;
  g_run = Schedule();
}

}  // namespace quijibo
