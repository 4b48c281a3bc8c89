// Bound-constrained limited-memory variable metric minimisation: the TOUPEE
// solver the reference's optimizer drivers call (TAO_SOLVER::optimizeBLMVM,
// src/TOUPEE/TAO_SOLVER.cpp:72-341), with its matrix (BLMVM.cpp:105-1018:
// limited-memory BFGS two-loop solve over a Broyden-updated diagonal, scalar
// rescaling, 5 pairs) and its bounded More-Thuente line search
// (MORE_THUENTE.cpp:104-818, LINE_SEARCH.cpp:35-300). The objective is the
// caller's callback; in the shape optimizer that is the GPU gradient pass.
//
// The iteration follows the reference step for step (sums are sequential;
// the reference's Eigen reductions may group them differently, which changes
// results only by rounding), including its state
// that persists across line searches (the interval lower end, the bracket
// flag and the last step code) and its handling of the "sufficient decrease
// at the bound step" exit, which leaves the search without a success code so
// that the solver restarts from the previous point with a scaled gradient.
#include <math.h>
#include <stdio.h>

#include <algorithm>
#include <vector>

#include "../../include/quijibo/quijibo.hpp"

namespace quijibo {

namespace {

constexpr Real kZeroSafeguard = 1e-8;  // BLMVM.h:36

Real dot(const VECTOR& a, const VECTOR& b) {
  Real s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

Real norm2(const VECTOR& a) { return sqrt(dot(a, a)); }

// Gradient components that push a variable sitting on a bound further out
// are dropped (BLMVM.cpp:958-979).
void bound_projection(const VECTOR& g, const VECTOR& x, const VECTOR& xl, const VECTOR& xu, VECTOR& out) {
  out.resize(g.size());
  for (size_t i = 0; i < g.size(); i++) {
    Real v = g[i];
    if (v > 0 && x[i] <= xl[i]) v = 0;
    else if (v < 0 && x[i] >= xu[i]) v = 0;
    out[i] = v;
  }
}

// median(lo, x, hi) per component
void clamp_to(VECTOR& x, const VECTOR& lo, const VECTOR& hi) {
  for (size_t i = 0; i < x.size(); i++) {
    Real a = lo[i], b = x[i], c = hi[i];
    // the middle value of three
    Real m = std::max(std::min(a, b), std::min(std::max(a, b), c));
    x[i] = m;
  }
}

// Limited-memory inverse-Hessian approximation: up to kPairs (s, y) pairs
// applied by the two-loop recursion around a positive diagonal D. D starts at
// delta and is updated per accepted pair by a Broyden mix (phi) of the BFGS
// and DFP diagonal updates, inverted, and rescaled by ys / (y' D y).
class VariableMetric {
 public:
  static constexpr int kPairs = 5;

  void set_delta(Real d) {
    delta_ = fabs(d);
    delta_ = std::max(delta_, (Real)1e-7);
    delta_ = std::min(delta_, (Real)100.0);
  }
  Real delta() const { return delta_; }

  void reset() {
    pairs_ = 0;
    updates_ = 0;
    calls_ = 0;
    std::fill(D_.begin(), D_.end(), delta_);
  }

  // Record the move from the previous (x, g) to this one.
  void update(const VECTOR& x, const VECTOR& g) {
    const size_t n = x.size();
    if (D_.size() != n) {
      D_.assign(n, delta_);
      S_.assign(kPairs, VECTOR(n, 0.0));
      Y_.assign(kPairs, VECTOR(n, 0.0));
      rho_.assign(kPairs, 0.0);
      beta_.assign(kPairs, 0.0);
    }
    if (calls_ == 0) {
      reset();
    } else {
      VECTOR s(n), y(n);
      for (size_t i = 0; i < n; i++) {
        s[i] = x[i] - xprev_[i];
        y[i] = g[i] - gprev_[i];
      }
      Real ys = dot(y, s);
      if (ys > 0.0) {  // curvature pair accepted (tolerance 0 * y'y)
        updates_++;
        pairs_ = std::min(pairs_ + 1, kPairs);
        for (int k = kPairs - 1; k > 0; k--) {
          S_[k].swap(S_[k - 1]);
          Y_[k].swap(Y_[k - 1]);
          rho_[k] = rho_[k - 1];
        }
        S_[0] = s;
        Y_[0] = y;
        rho_[0] = 1.0 / ys;
        update_diagonal(s, y, ys);
      }
    }
    calls_++;
    xprev_ = x;
    gprev_ = g;
  }

  // out = H b
  void solve(const VECTOR& b, VECTOR& out) {
    out = b;
    for (int k = 0; k < pairs_; k++) {
      beta_[k] = rho_[k] * dot(out, S_[k]);
      for (size_t i = 0; i < out.size(); i++) out[i] -= beta_[k] * Y_[k][i];
    }
    for (size_t i = 0; i < out.size(); i++) out[i] *= D_[i];
    for (int k = pairs_ - 1; k >= 0; k--) {
      const Real c = beta_[k] - rho_[k] * dot(out, Y_[k]);
      for (size_t i = 0; i < out.size(); i++) out[i] += c * S_[k][i];
    }
  }

 private:
  void update_diagonal(const VECTOR& s, const VECTOR& y, Real ys) {
    const size_t n = s.size();
    const Real phi = 0.125;
    // inverse diagonal and s' D^-1 s
    VECTOR inv(n), Ds(n);
    Real sDs = 0;
    for (size_t i = 0; i < n; i++) {
      inv[i] = 1.0 / D_[i];
      Ds[i] = s[i] * inv[i];
      sDs += Ds[i] * s[i];
    }
    if (ys == 0.0) ys = kZeroSafeguard;
    if (sDs == 0.0) sDs = kZeroSafeguard;
    // scalar coefficients first, then per component, in the reference's order
    const Real a_bfgs = 1.0 / ys, b_bfgs = -1.0 / sDs;
    const Real a_dfp = 1.0 / ys + sDs / (ys * ys), b_dfp = -2.0 / ys;
    VECTOR U(n);
    for (size_t i = 0; i < n; i++) {
      const Real yy = y[i] * y[i];
      // BFGS diagonal of the Hessian approximation
      const Real bfgs = (inv[i] + a_bfgs * yy) + b_bfgs * (Ds[i] * Ds[i]);
      // DFP diagonal
      const Real dfp = (inv[i] + a_dfp * yy) + b_dfp * (Ds[i] * y[i]);
      const Real mix = (1.0 - phi) * bfgs + phi * dfp;
      U[i] = fabs(1.0 / mix);
    }
    // scalar rescaling with the newest pair against the new diagonal
    Real yy_sum = 0, ys_sum = ys;
    for (size_t i = 0; i < n; i++) yy_sum += y[i] * U[i] * y[i];
    if (yy_sum == 0.0) ys_sum = kZeroSafeguard;  // the reference safeguards ys here (BLMVM.cpp:634-640)
    const Real sigma = ys_sum / yy_sum;
    if (!(isnan(sigma) || isinf(sigma)) && sigma != 0.0)
      for (size_t i = 0; i < n; i++) U[i] *= sigma;
    D_ = U;
  }

  Real delta_ = 1.0;
  int pairs_ = 0, updates_ = 0, calls_ = 0;
  VECTOR D_, xprev_, gprev_;
  std::vector<VECTOR> S_, Y_;
  std::vector<Real> rho_, beta_;
};

enum LsReason {
  LS_CONTINUE = 0,
  LS_SUCCESS = 1,
  LS_FAILED_ASCENT = -1,
  LS_FAILED_INFNAN = -2,
  LS_HALTED_OTHER = 2,
  LS_HALTED_UPPER = 3,
  LS_HALTED_LOWER = 4,
  LS_HALTED_RTOL = 5,
};

// Bounded More-Thuente search for a step meeting the strong Wolfe conditions
// (sufficient decrease ftol = 1e-4, curvature gtol = 0.9) within 30 trials.
// One object serves all searches of a solve; like the reference, it keeps
// the interval's lower end, the bracket flag and the last step code between
// searches.
class MoreThuente {
 public:
  explicit MoreThuente(FunctionGradient fg) : fg_(fg) {}

  LsReason search(VECTOR& x, Real& f, VECTOR& g, VECTOR& s, const VECTOR& xl, const VECTOR& xu) {
    const Real kInf = 1.0e100, xtrapf = 4.0;
    const size_t n = x.size();
    // keep the direction inside the box
    for (size_t i = 0; i < n; i++) s[i] = -s[i];
    bound_projection(s, x, xl, xu, s);
    for (size_t i = 0; i < n; i++) s[i] = -s[i];
    Real to_bound_min, wolfe_min, to_bound_max;
    step_bounds(x, xl, xu, s, to_bound_min, wolfe_min, to_bound_max);
    hi_ = std::min(to_bound_max, (Real)1.0e15);
    const Real dginit = dot(g, s);
    if (isnan(dginit) || isinf(dginit)) return LS_FAILED_INFNAN;
    if (dginit >= 0.0) return LS_FAILED_ASCENT;
    bracket_ = false;
    bool stage1 = true;
    const Real finit = f, dgtest = ftol_ * dginit;
    Real width = hi_ - lo_, width1 = 2.0 * width;
    Real stx = 0.0, fx = finit, dgx = dginit;
    Real sty = 0.0, fy = finit, dgy = dginit;
    step_ = 1.0;
    VECTOR trial(n);
    LsReason reason = LS_CONTINUE;
    for (int it = 0; it < 30; it++) {
      // interval of uncertainty for this trial
      if (bracket_) {
        lo_ = std::min(stx, sty);
        hi_ = std::max(stx, sty);
      } else {
        lo_ = stx;
        hi_ = step_ + xtrapf * (step_ - stx);
      }
      step_ = std::min(std::max(step_, lo_), hi_);
      if (stx != 0 && ((bracket_ && (step_ <= lo_ || step_ >= hi_)) || (bracket_ && hi_ - lo_ <= rtol_ * hi_) ||
                       code_ == 0))
        step_ = stx;
      for (size_t i = 0; i < n; i++) trial[i] = x[i] + step_ * s[i];
      clamp_to(trial, xl, xu);
      fg_(trial, f, g);
      Real dg = (dot(g, trial) - dot(g, x)) / step_;
      if (isinf(f) || isnan(f) || isinf(dg) || isnan(dg)) {
        f = kInf;
        dg = kInf;
      }
      const Real ftest1 = finit + step_ * dgtest;
      const Real ftest2 = finit + step_ * dgtest * ftol_;
      if (f - ftest1 <= 1.0e-10 * fabs(finit) && fabs(dg) + gtol_ * dginit <= 0.0) {
        reason = LS_SUCCESS;
        break;
      }
      // sufficient decrease at or past the nearest bound: stop with no code
      if (f <= ftest2 && step_ >= wolfe_min) break;
      if ((bracket_ && (step_ <= lo_ || step_ >= hi_)) || code_ == 0) {
        reason = LS_HALTED_OTHER;
        break;
      }
      if (step_ == hi_ && f <= ftest1 && dg <= dgtest) {
        reason = LS_HALTED_UPPER;
        break;
      }
      if (step_ == lo_ && f >= ftest1 && dg >= dgtest) {
        reason = LS_HALTED_LOWER;
        break;
      }
      if (bracket_ && hi_ - lo_ <= rtol_ * hi_) {
        reason = LS_HALTED_RTOL;
        break;
      }
      if (stage1 && f <= ftest1 && dg >= dginit * std::min(ftol_, gtol_)) stage1 = false;
      if (stage1 && f <= fx && f > ftest1) {
        // work on the auxiliary function f(step) - step * dgtest until it is
        // sufficiently decreased
        Real fxm = fx - stx * dgtest, fym = fy - sty * dgtest, fm = f - step_ * dgtest;
        Real dgxm = dgx - dgtest, dgym = dgy - dgtest, dgm = dg - dgtest;
        safeguarded_step(stx, fxm, dgxm, sty, fym, dgym, fm, dgm);
        fx = fxm + stx * dgtest;
        fy = fym + sty * dgtest;
        dgx = dgxm + dgtest;
        dgy = dgym + dgtest;
      } else {
        safeguarded_step(stx, fx, dgx, sty, fy, dgy, f, dg);
      }
      if (bracket_) {
        if (fabs(sty - stx) >= 0.66 * width1) step_ = stx + 0.5 * (sty - stx);
        width1 = width;
        width = fabs(sty - stx);
      }
    }
    x = trial;  // the last trial point, as the reference leaves it
    return reason;
  }

 private:
  // Distances along s to the box (LINE_SEARCH.cpp:70-136): the nearest bound,
  // the nearest bound among those ahead while all are ahead, and the farthest.
  static void step_bounds(const VECTOR& x, const VECTOR& xl, const VECTOR& xu, const VECTOR& s, Real& nearest,
                          Real& nearest_ahead, Real& farthest) {
    Real mn = 1.0e300, mw = 1.0e300, mx = 0.0;
    for (size_t i = 0; i < x.size(); i++) {
      Real t;
      if (s[i] > 0) t = (xu[i] - x[i]) / s[i];
      else if (s[i] < 0) t = (xl[i] - x[i]) / s[i];
      else continue;
      mn = std::min(t, mn);
      if (mn > 0) mw = std::min(t, mw);
      mx = std::max(t, mx);
    }
    nearest = mn;
    nearest_ahead = mw;
    farthest = mx;
  }

  // One safeguarded interpolation step (More and Thuente 1994, the four
  // cases of their step rule): updates the interval (stx, sty) from the trial
  // (step_, fp, dp) and chooses the next trial inside [lo_, hi_].
  void safeguarded_step(Real& stx, Real& fx, Real& dx, Real& sty, Real& fy, Real& dy, Real fp, Real dp) {
    const Real stp = step_;
    code_ = 0;
    if (bracket_ && (stp <= std::min(stx, sty) || stp >= std::max(stx, sty))) return;
    if (dx * (stp - stx) >= 0.0) return;
    if (hi_ < lo_) return;
    const Real sgnd = dp * (dx / fabs(dx));
    // minimiser of the cubic through (a, fa, da), (b, fb, db), measured from a;
    // `clip` keeps the square root's argument non-negative
    auto cubic = [](Real a, Real fa, Real da, Real b, Real fb, Real db, bool flip_when_b_below, bool clip) {
      const Real theta = 3 * (fa - fb) / (b - a) + da + db;
      Real sc = std::max(fabs(theta), fabs(da));
      sc = std::max(sc, fabs(db));
      Real rad = pow(theta / sc, (Real)2.0) - (da / sc) * (db / sc);
      if (clip) rad = std::max((Real)0.0, rad);
      Real gamma = sc * sqrt(rad);
      if (flip_when_b_below ? (b < a) : (b > a)) gamma = -gamma;
      return std::make_pair(theta, gamma);
    };
    Real stpf;
    bool bound;
    if (fp > fx) {
      // 1: higher value: the minimum is bracketed; cubic step if closer to stx
      code_ = 1;
      bound = true;
      auto tg = cubic(stx, fx, dx, stp, fp, dp, true, false);
      const Real p = (tg.second - dx) + tg.first, q = ((tg.second - dx) + tg.second) + dp;
      const Real stpc = stx + (p / q) * (stp - stx);
      const Real stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) * 0.5) * (stp - stx);
      stpf = fabs(stpc - stx) < fabs(stpq - stx) ? stpc : stpc + 0.5 * (stpq - stpc);
      bracket_ = true;
    } else if (sgnd < 0.0) {
      // 2: lower value, derivative changed sign: bracketed; the step farther from stp
      code_ = 2;
      bound = false;
      auto tg = cubic(stx, fx, dx, stp, fp, dp, false, false);
      const Real p = (tg.second - dp) + tg.first, q = ((tg.second - dp) + tg.second) + dx;
      const Real stpc = stp + (p / q) * (stx - stp);
      const Real stpq = stp + (dp / (dp - dx)) * (stx - stp);
      stpf = fabs(stpc - stp) > fabs(stpq - stp) ? stpc : stpq;
      bracket_ = true;
    } else if (fabs(dp) < fabs(dx)) {
      // 3: lower value, same sign, derivative shrinking
      code_ = 3;
      bound = true;
      auto tg = cubic(stx, fx, dx, stp, fp, dp, false, true);
      const Real p = (tg.second - dp) + tg.first, q = (tg.second + (dx - dp)) + tg.second;
      const Real r = p / q;
      Real stpc;
      if (r < 0.0 && tg.second != 0.0) stpc = stp + r * (stx - stp);
      else if (stp > stx) stpc = hi_;
      else stpc = lo_;
      const Real stpq = stp + (dp / (dp - dx)) * (stx - stp);
      if (bracket_) stpf = fabs(stp - stpc) < fabs(stp - stpq) ? stpc : stpq;
      else stpf = fabs(stp - stpc) > fabs(stp - stpq) ? stpc : stpq;
    } else {
      // 4: lower value, same sign, derivative not shrinking
      code_ = 4;
      bound = false;
      if (bracket_) {
        // cubic through the trial and the far end sty, measured from stp
        const Real theta = 3 * (fp - fy) / (sty - stp) + dy + dp;
        Real sc = std::max(std::max(fabs(theta), fabs(dy)), fabs(dp));
        Real gamma = sc * sqrt(pow(theta / sc, (Real)2.0) - (dy / sc) * (dp / sc));
        if (stp > sty) gamma = -gamma;
        const Real p = (gamma - dp) + theta, q = ((gamma - dp) + gamma) + dy;
        stpf = stp + (p / q) * (sty - stp);
      } else {
        stpf = stp > stx ? hi_ : lo_;
      }
    }
    // new interval
    if (fp > fx) {
      sty = stp;
      fy = fp;
      dy = dp;
    } else {
      if (sgnd < 0.0) {
        sty = stx;
        fy = fx;
        dy = dx;
      }
      stx = stp;
      fx = fp;
      dx = dp;
    }
    stpf = std::max(lo_, std::min(hi_, stpf));
    step_ = stpf;
    if (bracket_ && bound) {
      if (sty > stx) step_ = std::min(stx + (Real)0.66 * (sty - stx), step_);
      else step_ = std::max(stx + (Real)0.66 * (sty - stx), step_);
    }
  }

  FunctionGradient fg_;
  const Real ftol_ = 1e-4, gtol_ = 0.9, rtol_ = 1e-10;
  Real lo_ = 1e-20, hi_ = 1e20, step_ = 1.0;
  bool bracket_ = false;
  int code_ = 1;
};

}  // namespace

void TAO_SOLVER::SetTolerances(Real fatol, Real frtol, Real gatol, Real grtol, Real gttol) {
  if (fatol >= 0) _fatol = fatol;
  if (frtol >= 0) _frtol = frtol;
  if (gatol >= 0) _gatol = gatol;
  if (grtol >= 0) _grtol = grtol;
  if (gttol >= 0) _gttol = gttol;
}

int TAO_SOLVER::optimizeBLMVM(VECTOR& x, const VECTOR& xu, const VECTOR& xl, FunctionGradient fg, Monitor monitor) {
  const size_t n = x.size();
  if (xl.size() != n || xu.size() != n) throw Error(QJ_ERR_INVALID, "optimizeBLMVM: bounds must match the state");
  _history.clear();
  clamp_to(x, xl, xu);
  MoreThuente ls(fg);
  VariableMetric M;
  VECTOR g_raw(n), g(n), dir(n), x_old, g_old;
  Real f = 0;
  fg(x, f, g_raw);
  bound_projection(g_raw, x, xl, xu, g);
  Real gnorm = norm2(g);
  _gnorm0 = gnorm;
  auto scaled_delta = [&]() { return f != 0.0 ? 2.0 * fabs(f) / (gnorm * gnorm) : 2.0 / (gnorm * gnorm); };
  M.set_delta(scaled_delta());
  _niter = 0;
  _reason = CONTINUE_ITERATING;
  _history.push_back(f);
  while (_reason == CONTINUE_ITERATING) {
    M.update(x, g);
    M.solve(g_raw, dir);
    bound_projection(dir, x, xl, xu, g);  // g holds the projected direction until the step is taken
    if (dot(g_raw, g) <= 0) {
      // not a descent direction: restart from the scaled gradient
      M.set_delta(scaled_delta());
      M.reset();
      M.update(x, g_raw);
      M.solve(g_raw, dir);
    }
    for (Real& v : dir) v = -v;
    const Real f_old = f;
    x_old = x;
    g_old = g_raw;
    LsReason r = ls.search(x, f, g_raw, dir, xl, xu);
    if (r != LS_SUCCESS) {
      // restart from the previous point along the scaled gradient
      f = f_old;
      x = x_old;
      g_raw = g_old;
      M.set_delta(scaled_delta());
      M.reset();
      M.update(x, g_raw);
      M.solve(g_raw, dir);
      for (Real& v : dir) v = -v;
      r = ls.search(x, f, g_raw, dir, xl, xu);
      if (r != LS_SUCCESS) {
        _reason = DIVERGED_LS_FAILURE;
        break;
      }
    }
    bound_projection(g_raw, x, xl, xu, g);
    gnorm = norm2(g);
    _niter++;
    _history.push_back(f);
    converged(f, gnorm);
    if (monitor) monitor();
  }
  _f = f;
  return _niter;
}

// TAO's default convergence test (TAO_SOLVER.cpp:695-830) with no constraint
// norms and no function-evaluation count.
void TAO_SOLVER::converged(Real f, Real gnorm) {
  const Real g2 = gnorm * gnorm;
  if (isinf(f) || isnan(f)) _reason = DIVERGED_NAN;
  else if (f <= _fmin) _reason = CONVERGED_MINF;
  else if (g2 <= _fatol) _reason = CONVERGED_FATOL;
  else if (f != 0 && g2 / fabs(f) <= _frtol) _reason = CONVERGED_FRTOL;
  else if (gnorm <= _gatol) _reason = CONVERGED_GATOL;
  else if (f != 0 && fabs(gnorm / f) <= _grtol) _reason = CONVERGED_GRTOL;
  else if (_gnorm0 != 0 && gnorm / _gnorm0 <= _gttol) _reason = CONVERGED_GTTOL;
  else if (_niter > _maxIterations) _reason = DIVERGED_MAXITS;
  else _reason = CONTINUE_ITERATING;
}

}  // namespace quijibo
