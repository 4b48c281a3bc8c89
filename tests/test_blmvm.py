"""The TOUPEE BLMVM solver: the oracle restatement (oracle/blmvm.py) is pinned
by closed-form and independent-solver optima, and the product's C++ solver
(quijibo::TAO_SOLVER, run through tests/cpp/host_layer_driver) follows the
oracle iterate by iterate."""
import json
import os
import subprocess

import numpy as np
import pytest
import scipy.optimize as so

from oracle import blmvm as B

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "host_layer_driver")
FLT_MAX = 3.4028234663852886e38


def rosenbrock(x):
    # the same operation order as the driver's C++ function
    n = len(x)
    f = 0.0
    g = [0.0] * n
    for i in range(n - 1):
        a = x[i + 1] - x[i] * x[i]
        b = 1.0 - x[i]
        f += 100.0 * a * a + b * b
        g[i] += -400.0 * a * x[i] - 2.0 * b
        g[i + 1] += 200.0 * a
    return f, g


BQ_B = [1.0, -2.0, 3.0, 0.5, -1.0, 2.0]


def boxquad(x):
    n = len(x)
    f = 0.0
    g = [0.0] * n
    for i in range(n):
        ax = 2.5 * x[i]
        if i > 0:
            ax -= x[i - 1]
        if i + 1 < n:
            ax -= x[i + 1]
        g[i] = ax - BQ_B[i]
        f += 0.5 * x[i] * ax - BQ_B[i] * x[i]
    return f, g


PROBLEMS = {
    "rosenbrock": (rosenbrock, [-1.2, 1.0, -1.2, 1.0], [-FLT_MAX] * 4, [FLT_MAX] * 4),
    "boxquad": (boxquad, [0.0] * 6, [-0.5] * 6, [0.8] * 6),
}


def test_oracle_rosenbrock_reaches_the_minimum():
    fg, x0, lo, hi = PROBLEMS["rosenbrock"]
    x, hist, it, reason = B.optimize_blmvm(x0, hi, lo, fg)
    assert reason > 0 and np.allclose(x, 1.0, atol=1e-5) and hist[-1] < 1e-10
    assert all(b <= a for a, b in zip(hist, hist[1:]))  # every accepted step decreases f


def test_oracle_bounded_quadratic_matches_an_independent_solver():
    fg, x0, lo, hi = PROBLEMS["boxquad"]
    x, hist, it, reason = B.optimize_blmvm(x0, hi, lo, fg)
    ref = so.minimize(lambda v: fg(list(v))[0], np.zeros(6), jac=lambda v: np.array(fg(list(v))[1]),
                      method="L-BFGS-B", bounds=[(-0.5, 0.8)] * 6, options={"ftol": 1e-15, "gtol": 1e-12})
    # TAO stops at ||projected g||^2 <= fatol = 1e-8, i.e. ||g|| <= 1e-4; A's smallest
    # eigenvalue is 0.77, so x is within ~1.3e-4 of the optimum and f within ~1e-8
    assert np.allclose(x, ref.x, atol=2e-4) and abs(hist[-1] - ref.fun) < 1e-8
    # KKT: the projected gradient vanishes
    g = np.array(fg(x)[1])
    xa = np.array(x)
    pg = np.where((g > 0) & (xa <= -0.5), 0.0, np.where((g < 0) & (xa >= 0.8), 0.0, g))
    assert np.max(np.abs(pg)) < 1e-4
    assert xa[2] == 0.8 and xa[5] == 0.8  # active upper bounds


def test_oracle_respects_bounds_from_an_infeasible_start():
    fg, _, lo, hi = PROBLEMS["boxquad"]
    x, hist, it, reason = B.optimize_blmvm([3.0, -3.0, 3.0, -3.0, 3.0, -3.0], hi, lo, fg)
    assert all(-0.5 <= v <= 0.8 for v in x) and reason > 0


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="host_layer_driver not built")
@pytest.mark.parametrize("problem", ["rosenbrock", "boxquad"])
def test_cpp_solver_follows_the_oracle(problem):
    r = subprocess.run([DRIVER, "blmvm", problem], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    fg, x0, lo, hi = PROBLEMS[problem]
    x, hist, it, reason = B.optimize_blmvm(x0, hi, lo, fg)
    assert got["iterations"] == it and got["reason"] == reason
    assert np.allclose(got["history"], hist, rtol=1e-12, atol=1e-14)
    assert np.allclose(got["x"], x, rtol=1e-12, atol=1e-14)
