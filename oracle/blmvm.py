"""CPU restatement of the reference's bound-constrained optimizer (test infrastructure).

TAO_SOLVER::optimizeBLMVM (src/TOUPEE/TAO_SOLVER.cpp:72-341) with its matrix
(BLMVM.cpp: MatReset :105-147, Update :176-830 in the configuration the solver
uses -- Broyden diagonal scaling with phi = 0.125, scalar rescaling with
r_alpha = 1, r_beta = 0.5, one history entry, 5 pairs -- MatSolve :832-956,
VecBoundGradientProjection :958-979, SetDelta :1014-1019), its line search
(MORE_THUENTE.cpp: mcstep :104-408, LineSearchImplementation :410-818;
LINE_SEARCH.cpp: defaults :35-58, VecStepBoundInfo :70-136) and its
convergence test (TAO_SOLVER.cpp:695-830).

Written in plain Python with sequential sums so that every operation happens
in the same order as in the reference; only tests import it, to check the
product's C++ solver iterate by iterate.
"""
import math

ZERO_SAFEGUARD = 1e-8  # BLMVM.h:36


def _dot(a, b):
    s = 0.0
    for x, y in zip(a, b):
        s += x * y
    return s


def _projection(g, x, xl, xu):
    """VecBoundGradientProjection (BLMVM.cpp:958-979)."""
    out = []
    for gi, xi, li, ui in zip(g, x, xl, xu):
        v = gi
        if v > 0 and xi <= li:
            v = 0.0
        elif v < 0 and xi >= ui:
            v = 0.0
        out.append(v)
    return out


def _median(lo, x, hi):
    return [sorted((a, b, c))[1] for a, b, c in zip(lo, x, hi)]


class _Matrix:
    """BLMVM's LMVM matrix (lm = 5, Broyden diagonal, scalar rescale)."""

    def __init__(self):
        self.lm = 5
        self.delta = 1.0
        self.S, self.Y = [], []
        self.rho = [0.0] * 6
        self.beta = [0.0] * 6
        self.D = None
        self.iter = 0
        self.lmnow = 0
        self.nupdates = 0
        self.xprev = self.gprev = None

    def set_delta(self, d):  # SetDelta, BLMVM.cpp:1014-1019
        d = abs(d)
        d = d if d > 1e-7 else 1e-7
        self.delta = d if d < 100 else 100.0

    def reset(self, n):  # MatReset, BLMVM.cpp:105-147
        self.rho = [0.0] * 6
        self.rho[0] = 1.0
        self.D = [self.delta] * n
        self.iter = 0
        self.nupdates = 0
        self.lmnow = 0

    def update(self, x, g):  # Update, BLMVM.cpp:176-830
        n = len(x)
        if not self.S:
            self.S = [[0.0] * n for _ in range(6)]
            self.Y = [[0.0] * n for _ in range(6)]
        if self.iter == 0:
            self.reset(n)
        else:
            y = [gi - pi for gi, pi in zip(g, self.gprev)]
            s = [xi - pi for xi, pi in zip(x, self.xprev)]
            rhotemp = _dot(y, s)
            y0temp = _dot(y, y)
            if rhotemp > 0.0 * y0temp:
                self.nupdates += 1
                self.lmnow = min(self.lmnow + 1, self.lm)
                for i in range(self.lm - 1, -1, -1):
                    self.S[i + 1] = self.S[i]
                    self.Y[i + 1] = self.Y[i]
                    self.rho[i + 1] = self.rho[i]
                self.S[0] = s
                self.Y[0] = y
                self.rho[0] = 1.0 / rhotemp
                # Broyden diagonal (:387-470)
                P = [1.0 / d for d in self.D]
                Q = list(P)
                V = [yi * yi for yi in y]
                W = [si * pi for si, pi in zip(s, P)]
                sDs = _dot(W, s)
                if rhotemp == 0.0:
                    rhotemp = ZERO_SAFEGUARD
                if sDs == 0.0:
                    sDs = ZERO_SAFEGUARD
                phi = 0.125
                U = [wi * wi for wi in W]
                P = [p + 1.0 / rhotemp * v for p, v in zip(P, V)]
                P = [p + -1.0 / sDs * u for p, u in zip(P, U)]
                U = [wi * yi for wi, yi in zip(W, y)]
                Q = [q + (1.0 / rhotemp + sDs / (rhotemp * rhotemp)) * v for q, v in zip(Q, V)]
                Q = [q + -2.0 / rhotemp * u for q, u in zip(Q, U)]
                U = [(1.0 - phi) * p + phi * q for p, q in zip(P, Q)]
                U = [abs(1.0 / u) for u in U]
                # scalar rescale with the current diagonal (:473-640)
                P = list(U)
                V = [yi * pi for yi, pi in zip(self.Y[0], P)]
                yy_sum = _dot(V, self.Y[0])
                ys_sum = rhotemp
                if yy_sum == 0.0:
                    ys_sum = ZERO_SAFEGUARD
                sigmanew = ys_sum / yy_sum if yy_sum != 0.0 else math.inf
                if not (math.isnan(sigmanew) or math.isinf(sigmanew)) and sigmanew != 0.0:
                    U = [u * sigmanew for u in U]
                self.D = U
        self.iter += 1
        self.xprev = list(x)
        self.gprev = list(g)

    def solve(self, b):  # MatSolve, BLMVM.cpp:832-956
        x = list(b)
        for ll in range(self.lmnow):
            sq = _dot(x, self.S[ll])
            self.beta[ll] = sq * self.rho[ll]
            x = [xi + -self.beta[ll] * yi for xi, yi in zip(x, self.Y[ll])]
        x = [xi * di for xi, di in zip(x, self.D)]
        for ll in range(self.lmnow - 1, -1, -1):
            yq = _dot(x, self.Y[ll])
            c = self.beta[ll] - yq * self.rho[ll]
            x = [xi + c * si for xi, si in zip(x, self.S[ll])]
        return x


SUCCESS, CONTINUE = 1, 0


class _MoreThuente:
    """MORE_THUENTE with the LINE_SEARCH defaults; state persists across searches."""

    def __init__(self, fg):
        self.fg = fg
        self.ftol, self.gtol, self.rtol = 1e-4, 0.9, 1e-10
        self.stepmin, self.stepmax, self.step = 1e-20, 1e20, 1.0
        self.bracket = 0
        self.infoc = 1
        self.max_funcs = 30

    def mcstep(self, stx, fx, dx, sty, fy, dy, stp, fp, dp):  # :104-408
        self.infoc = 0
        if self.bracket and (stp <= min(stx, sty) or stp >= max(stx, sty)):
            return stx, fx, dx, sty, fy, dy, stp
        if dx * (stp - stx) >= 0.0:
            return stx, fx, dx, sty, fy, dy, stp
        if self.stepmax < self.stepmin:
            return stx, fx, dx, sty, fy, dy, stp
        sgnd = dp * (dx / abs(dx))
        if fp > fx:
            self.infoc, bound = 1, 1
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp
            s = max(abs(theta), abs(dx), abs(dp))
            gamma1 = s * math.sqrt((theta / s) ** 2.0 - (dx / s) * (dp / s))
            if stp < stx:
                gamma1 = -gamma1
            p = (gamma1 - dx) + theta
            q = ((gamma1 - dx) + gamma1) + dp
            r = p / q
            stpc = stx + r * (stp - stx)
            stpq = stx + ((dx / ((fx - fp) / (stp - stx) + dx)) * 0.5) * (stp - stx)
            stpf = stpc if abs(stpc - stx) < abs(stpq - stx) else stpc + 0.5 * (stpq - stpc)
            self.bracket = 1
        elif sgnd < 0.0:
            self.infoc, bound = 2, 0
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp
            s = max(abs(theta), abs(dx), abs(dp))
            gamma1 = s * math.sqrt((theta / s) ** 2.0 - (dx / s) * (dp / s))
            if stp > stx:
                gamma1 = -gamma1
            p = (gamma1 - dp) + theta
            q = ((gamma1 - dp) + gamma1) + dx
            r = p / q
            stpc = stp + r * (stx - stp)
            stpq = stp + (dp / (dp - dx)) * (stx - stp)
            stpf = stpc if abs(stpc - stp) > abs(stpq - stp) else stpq
            self.bracket = 1
        elif abs(dp) < abs(dx):
            self.infoc, bound = 3, 1
            theta = 3 * (fx - fp) / (stp - stx) + dx + dp
            s = max(abs(theta), abs(dx), abs(dp))
            gamma1 = s * math.sqrt(max(0.0, (theta / s) ** 2.0 - (dx / s) * (dp / s)))
            if stp > stx:
                gamma1 = -gamma1
            p = (gamma1 - dp) + theta
            q = (gamma1 + (dx - dp)) + gamma1
            r = p / q
            if r < 0.0 and gamma1 != 0.0:
                stpc = stp + r * (stx - stp)
            elif stp > stx:
                stpc = self.stepmax
            else:
                stpc = self.stepmin
            stpq = stp + (dp / (dp - dx)) * (stx - stp)
            if self.bracket:
                stpf = stpc if abs(stp - stpc) < abs(stp - stpq) else stpq
            else:
                stpf = stpc if abs(stp - stpc) > abs(stp - stpq) else stpq
        else:
            self.infoc, bound = 4, 0
            if self.bracket:
                theta = 3 * (fp - fy) / (sty - stp) + dy + dp
                s = max(abs(theta), abs(dy), abs(dp))
                gamma1 = s * math.sqrt((theta / s) ** 2.0 - (dy / s) * (dp / s))
                if stp > sty:
                    gamma1 = -gamma1
                p = (gamma1 - dp) + theta
                q = ((gamma1 - dp) + gamma1) + dy
                r = p / q
                stpf = stp + r * (sty - stp)
            elif stp > stx:
                stpf = self.stepmax
            else:
                stpf = self.stepmin
        if fp > fx:
            sty, fy, dy = stp, fp, dp
        else:
            if sgnd < 0.0:
                sty, fy, dy = stx, fx, dx
            stx, fx, dx = stp, fp, dp
        stpf = min(self.stepmax, stpf)
        stpf = max(self.stepmin, stpf)
        stp = stpf
        if self.bracket and bound:
            if sty > stx:
                stp = min(stx + 0.66 * (sty - stx), stp)
            else:
                stp = max(stx + 0.66 * (sty - stx), stp)
        return stx, fx, dx, sty, fy, dy, stp

    def search(self, x, f, g, s, xl, xu):  # :410-818 (bounded)
        n = len(x)
        s = [-v for v in s]
        s = _projection(s, x, xl, xu)
        s = [-v for v in s]
        bmin, wolfemin, bmax = 1e300, 1e300, 0.0  # VecStepBoundInfo
        for i in range(n):
            if s[i] > 0:
                t = (xu[i] - x[i]) / s[i]
            elif s[i] < 0:
                t = (xl[i] - x[i]) / s[i]
            else:
                continue
            bmin = min(t, bmin)
            if bmin > 0:
                wolfemin = min(t, wolfemin)
            bmax = max(t, bmax)
        self.stepmax = min(bmax, 1.0e15)
        dginit = _dot(g, s)
        if math.isnan(dginit) or math.isinf(dginit) or dginit >= 0.0:
            return x, f, g, -1
        self.bracket = 0
        stage1 = 1
        finit = f
        dgtest = self.ftol * dginit
        width = self.stepmax - self.stepmin
        width1 = width * 2.0
        stx, fx, dgx = 0.0, finit, dginit
        sty, fy, dgy = 0.0, finit, dginit
        self.step = 1.0
        reason = CONTINUE
        work = list(x)
        for _ in range(self.max_funcs):
            if self.bracket:
                self.stepmin = min(stx, sty)
                self.stepmax = max(stx, sty)
            else:
                self.stepmin = stx
                self.stepmax = self.step + 4.0 * (self.step - stx)
            self.step = max(self.step, self.stepmin)
            self.step = min(self.step, self.stepmax)
            if stx != 0 and ((self.bracket and (self.step <= self.stepmin or self.step >= self.stepmax)) or
                             (self.bracket and (self.stepmax - self.stepmin <= self.rtol * self.stepmax)) or
                             self.infoc == 0):
                self.step = stx
            work = [xi + self.step * si for xi, si in zip(x, s)]
            work = _median(xl, work, xu)
            f, g = self.fg(work)
            dg = (_dot(g, work) - _dot(g, x)) / self.step
            if math.isinf(f) or math.isnan(f) or math.isinf(dg) or math.isnan(dg):
                f, dg = 1.0e100, 1.0e100
            ftest1 = finit + self.step * dgtest
            ftest2 = finit + self.step * dgtest * self.ftol
            if f - ftest1 <= 1.0e-10 * abs(finit) and abs(dg) + self.gtol * dginit <= 0.0:
                reason = SUCCESS
                break
            if f <= ftest2 and self.step >= wolfemin:
                break
            if (self.bracket and (self.step <= self.stepmin or self.step >= self.stepmax)) or not self.infoc:
                reason = 2
                break
            if self.step == self.stepmax and f <= ftest1 and dg <= dgtest:
                reason = 3
                break
            if self.step == self.stepmin and f >= ftest1 and dg >= dgtest:
                reason = 4
                break
            if self.bracket and self.stepmax - self.stepmin <= self.rtol * self.stepmax:
                reason = 5
                break
            if stage1 and f <= ftest1 and dg >= dginit * min(self.ftol, self.gtol):
                stage1 = 0
            if stage1 and f <= fx and f > ftest1:
                fm = f - self.step * dgtest
                fxm = fx - stx * dgtest
                fym = fy - sty * dgtest
                dgm = dg - dgtest
                dgxm = dgx - dgtest
                dgym = dgy - dgtest
                stx, fxm, dgxm, sty, fym, dgym, self.step = self.mcstep(stx, fxm, dgxm, sty, fym, dgym, self.step,
                                                                        fm, dgm)
                fx = fxm + stx * dgtest
                fy = fym + sty * dgtest
                dgx = dgxm + dgtest
                dgy = dgym + dgtest
            else:
                stx, fx, dgx, sty, fy, dgy, self.step = self.mcstep(stx, fx, dgx, sty, fy, dgy, self.step, f, dg)
            if self.bracket:
                if abs(sty - stx) >= 0.66 * width1:
                    self.step = stx + 0.5 * (sty - stx)
                width1 = width
                width = abs(sty - stx)
        return work, f, g, reason


def optimize_blmvm(x, xu, xl, fg, max_it=2000, tolerances=(1e-8, 1e-8, 1e-8, 1e-8, 0.0)):
    """Returns (x, history of f, iterations, reason). fg(x) -> (f, g)."""
    fatol, frtol, gatol, grtol, gttol = tolerances
    fmin = -1e100
    x = _median(xl, [float(v) for v in x], xu)
    ls = _MoreThuente(fg)
    M = _Matrix()
    f, g_raw = fg(x)
    g = _projection(g_raw, x, xl, xu)
    gnorm = math.sqrt(_dot(g, g))
    gnorm0 = gnorm

    def scaled_delta():
        return 2.0 * abs(f) / (gnorm * gnorm) if f != 0.0 else 2.0 / (gnorm * gnorm)

    M.set_delta(scaled_delta())
    it = 0
    reason = 0
    hist = [f]
    while reason == 0:
        M.update(x, g)
        d = M.solve(g_raw)
        g = _projection(d, x, xl, xu)
        if _dot(g_raw, g) <= 0:
            M.set_delta(scaled_delta())
            M.reset(len(x))
            M.update(x, g_raw)
            d = M.solve(g_raw)
        d = [-v for v in d]
        fold, xold, gold = f, list(x), list(g_raw)
        x, f, g_raw, r = ls.search(x, f, g_raw, d, xl, xu)
        if r != SUCCESS:
            f, x, g_raw = fold, xold, gold
            M.set_delta(scaled_delta())
            M.reset(len(x))
            M.update(x, g_raw)
            d = [-v for v in M.solve(g_raw)]
            x, f, g_raw, r = ls.search(x, f, g_raw, d, xl, xu)
            if r != SUCCESS:
                reason = -6
                break
        g = _projection(g_raw, x, xl, xu)
        gnorm = math.sqrt(_dot(g, g))
        it += 1
        hist.append(f)
        g2 = gnorm * gnorm
        if math.isinf(f) or math.isnan(f):
            reason = -4
        elif f <= fmin:
            reason = 7
        elif g2 <= fatol:
            reason = 1
        elif f != 0 and g2 / abs(f) <= frtol:
            reason = 2
        elif gnorm <= gatol:
            reason = 3
        elif f != 0 and abs(gnorm / f) <= grtol:
            reason = 4
        elif gnorm0 != 0 and gnorm / gnorm0 <= gttol:
            reason = 5
        elif it > max_it:
            reason = -2
    return x, hist, it, reason
