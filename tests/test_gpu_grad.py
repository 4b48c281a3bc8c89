"""GPU gradients against the oracle and against finite differences.

Tolerances:
  AP score/gradients (FP64) vs oracle: relative 1e-9 on every entry (the
  kernel recovers prefixes as T S^-1 instead of caching them, a few ulp).
  Exact derivatives vs central FD (h = 1e-6): relative 1e-5 of the gradient norm.
  FP32 adjoint vs FP64 adjoint: relative 2e-3 of the gradient norm.
"""
import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api, grad as G, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_ADAPTIVE, MAP_BLACK_HOLE, MAP_SKIP_ORIGIN

pytestmark = pytest.mark.gpu


def _weights(res, lengths=(3.0, 3.0, 3.0)):
    # inverse-distance-like weight field on the same grid (OPTIMIZE_3D.cpp:10840-10843 uses 1/|d|)
    ax = [-lengths[i] / 2 + (np.arange(res[i]) + 0.5) * lengths[i] / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    d = np.abs(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.8) + 0.05
    return (1.0 / d).reshape(-1)


def _ap_both(w, res, wres=None, flags=MAP_SKIP_ORIGIN | MAP_ADAPTIVE, escape=20.0):
    wres = wres or res
    weight = _weights(wres)
    g = api.make_grid(res, w.center, w.lengths)
    wg = api.make_grid(wres, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers, w.top.power_scalar),
                     api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar))
    p = api.make_params(1, escape, w.exp_scaling, w.slice, flags)
    gpu = api.ap_score_gradients(g, r, p, wg, torch.from_numpy(weight).cuda())
    cpu = O.ap_score_gradients(O.make_grid(res, w.center, w.lengths),
                               O.poly(w.top.roots, w.top.powers, w.top.power_scalar),
                               O.poly(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar),
                               O.make_params(1, escape, w.exp_scaling, w.slice, O.SKIP_ORIGIN | O.AP_REPASS),
                               O.make_grid(wres, w.center, w.lengths), weight)
    return gpu, cpu


def _rel(a, b):
    a = np.atleast_1d(np.asarray(a, dtype=np.float64))
    b = np.atleast_1d(np.asarray(b, dtype=np.float64))
    return float(np.max(np.abs(a - b)) / max(1e-300, np.max(np.abs(b))))


def _check_ap(gpu, cpu, tol=1e-9):
    for k in ("score", "grad_exp", "grad_top_scalar", "grad_bottom_scalar", "grad_top", "grad_bottom"):
        e = _rel(gpu[k], cpu[k])
        print("%s rel err %.2e" % (k, e), flush=True)
        assert e <= tol, k
    f = gpu["field"].cpu().numpy()
    assert np.allclose(f, cpu["field"], rtol=1e-9, atol=1e-12)


def test_ap_gradients_config2_roots(gpu):
    w = W.config2(20)
    w.top.power_scalar, w.bottom.power_scalar, w.exp_scaling = 0.9, 1.2, 0.7
    gpu_r, cpu_r = _ap_both(w, (20, 18, 16), wres=(15, 17, 19))
    _check_ap(gpu_r, cpu_r)


def test_ap_gradients_overflow_cells(gpu):
    # T and B both overflow FP64 while T/B stays moderate, so the extended-range
    # cells carry non-zero gradient terms
    top = W.RootSet(np.array([[1.9, 0.3, 0.1, 0.0], [-2.1, 0.4, -0.2, 0.0], [0.2, -0.3, 0.1, 0.05]]),
                    np.array([200.0, 200.0, 1.0]))
    bot = W.RootSet(np.array([[1.902, 0.31, 0.1, 0.0], [-2.1, 0.41, -0.19, 0.0], [0.3, 0.2, -0.1, 0.0]]),
                    np.array([200.0, 200.0, 1.0]))
    w = W.Workload("ap", (12, 12, 11), top, bot, max_iter=1, escape=20.0, exp_scaling=0.5)
    gpu_r, cpu_r = _ap_both(w, (12, 12, 11))
    assert gpu_r["adaptive_cells"] > 0 and cpu_r["overflowed"] == gpu_r["adaptive_cells"]
    _check_ap(gpu_r, cpu_r)


def test_ap_top_gradient_is_the_score_derivative(gpu):
    """The top-power and exp gradients are exact derivatives of the score at one iteration."""
    w = W.config2(16)
    res = (16, 16, 16)
    weight = torch.from_numpy(_weights(res)).cuda()
    g = api.make_grid(res, w.center, w.lengths)
    p = api.make_params(1, 20.0, w.exp_scaling, 0.0, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
    top = api.POLYNOMIAL_4D(w.top.roots, w.top.powers)
    bot = api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers)
    r = api.Rational(top, bot)
    base = api.ap_score_gradients(g, r, p, g, weight)
    h = 1e-6
    for i in (0, 7, 31):
        vals = []
        for sgn in (1, -1):
            tp = top.powers.copy()
            tp[i] += sgn * h
            r.update(api.POLYNOMIAL_4D(w.top.roots, tp), bot)
            vals.append(api.ap_score_gradients(g, r, p, g, weight)["score"])
        fd = (vals[0] - vals[1]) / (2 * h)
        assert abs(fd - base["grad_top"][i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, base["grad_top"][i])
    r.update(top, bot)
    vals = []
    for sgn in (1, -1):
        pp = api.make_params(1, 20.0, w.exp_scaling * np.exp(sgn * h), 0.0, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
        vals.append(api.ap_score_gradients(g, r, pp, g, weight)["score"])
    fd = (vals[0] - vals[1]) / (2 * h)
    assert abs(fd - base["grad_exp"]) <= 1e-5 * max(1.0, abs(fd))


def _gpu_ap(case, tp=None, bp=None, st=None, sb=None):
    """The GPU AP pass on a test_oracle_grad case (same grid for the weights)."""
    g = api.make_grid(case["res"], case["center"], case["lengths"])
    r = api.Rational(api.POLYNOMIAL_4D(case["top"], case["tp"] if tp is None else tp, case["st"] if st is None else st),
                     api.POLYNOMIAL_4D(case["bot"], case["bp"] if bp is None else bp, case["sb"] if sb is None else sb))
    p = api.make_params(1, 20.0, case["c"], 0.0, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
    return api.ap_score_gradients(g, r, p, g, torch.from_numpy(case["weight"]).cuda(), 10.0, 0.0)


def test_ap_gradients_match_closed_form_on_commuting_plane(gpu):
    """All six outputs against the complex closed form (no oracle involved)."""
    from test_oracle_grad import plane_case, plane_closed_form
    case = plane_case()
    cf = plane_closed_form(case)
    r = _gpu_ap(case)
    assert r["adaptive_cells"] == 0
    for k in ("score", "grad_exp", "grad_top", "grad_bottom", "grad_top_scalar", "grad_bottom_scalar"):
        e = _rel(r[k], cf[k])
        print("%s rel err vs closed form %.2e" % (k, e), flush=True)
        assert e < 1e-10, k


def test_ap_scalar_and_bottom_gradients_are_score_derivatives(gpu):
    """grad_bottom, grad_top_scalar and grad_bottom_scalar against central
    differences of the GPU score on a general 3D grid (at one iteration every
    reference gradient formula is the exact derivative)."""
    rng = np.random.default_rng(12)
    case = dict(top=rng.normal(scale=0.6, size=(7, 4)), bot=rng.normal(scale=0.6, size=(6, 4)),
                tp=rng.uniform(0.8, 2.0, 7), bp=rng.uniform(0.8, 2.0, 6), st=0.95, sb=1.2, c=1.3,
                res=(20, 18, 16), center=(0.0, 0.05, 0.1), lengths=(2.6, 2.6, 2.6))
    case["weight"] = rng.uniform(0.5, 2.0, 20 * 18 * 16)
    base = _gpu_ap(case)
    assert base["adaptive_cells"] == 0
    h = 1e-6
    fdb = []
    for j in range(6):
        s = []
        for sg in (1, -1):
            bp = case["bp"].copy()
            bp[j] += sg * h
            s.append(_gpu_ap(case, bp=bp)["score"])
        fdb.append((s[0] - s[1]) / (2 * h))
    e = _rel(base["grad_bottom"], fdb)
    print("grad_bottom vs FD rel %.2e" % e, flush=True)
    assert e < 1e-6
    for key, kw in (("grad_top_scalar", "st"), ("grad_bottom_scalar", "sb")):
        s = [_gpu_ap(case, **{kw: case[kw] + sg * h})["score"] for sg in (1, -1)]
        fd = (s[0] - s[1]) / (2 * h)
        print("%s %.10e FD %.10e" % (key, base[key], fd), flush=True)
        assert abs(base[key] - fd) < 1e-6 * abs(fd), key


def _c5(n=20000, max_iter=3):
    w = W.config5(n_surface=n, n_exterior=n, max_iter=max_iter)
    pts = torch.from_numpy(np.concatenate([w.extra["surface"], w.extra["exterior"]])).cuda()
    return w, pts


def _objective(pts, w, r, p, n):
    F = api.points_eval(pts, r, p, FP64)
    return G.superquadric_objective(F, n)


def test_adjoint_matches_central_differences_full_c5(gpu):
    """Config 5: 20k surface + 20k exterior points, 64+64 roots, all 512 root
    components and 128 powers against central differences with h = 1e-6."""
    n = 20000
    w, pts = _c5(n)
    p = api.make_params(w.max_iter, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE)
    top = api.POLYNOMIAL_4D(w.top.roots.copy(), w.top.powers.copy())
    bot = api.POLYNOMIAL_4D(w.bottom.roots.copy(), w.bottom.powers.copy())
    r = api.Rational(top, bot)
    F = api.points_eval(pts, r, p, FP64)
    seed = G.superquadric_objective_seed(F, n)
    gr, gp = api.points_adjoint(pts, r, p, seed, FP64)
    gr32, gp32 = api.points_adjoint(pts, r, p, seed, FP32_BAND)
    fd = G.centered_root_gradient(lambda rr: _objective(pts, w, rr, p, n), top, bot, 1e-6)
    adj = gr.cpu().numpy().reshape(-1)
    scale = np.linalg.norm(fd)
    e64 = np.linalg.norm(adj - fd) / scale
    e32 = np.linalg.norm(gr32.cpu().numpy().reshape(-1) - adj) / np.linalg.norm(adj)
    print("adjoint fp64 vs FD rel %.2e, fp32 vs fp64 rel %.2e, |g| %.3e" % (e64, e32, scale), flush=True)
    assert e64 < 1e-5
    assert e32 < 2e-3
    # powers
    h = 1e-6
    fdp = []
    for which, poly in ((0, top), (1, bot)):
        for i in range(0, poly.totalRoots(), 9):
            v = []
            for sgn in (1, -1):
                old = poly.powers[i]
                poly.powers[i] = old + sgn * h
                r.update(top, bot)
                v.append(_objective(pts, w, r, p, n))
                poly.powers[i] = old
            r.update(top, bot)
            fdp.append(((v[0] - v[1]) / (2 * h), gp.cpu().numpy()[i + which * top.totalRoots()]))
    fdp = np.array(fdp)
    ep = np.linalg.norm(fdp[:, 0] - fdp[:, 1]) / np.linalg.norm(fdp[:, 0])
    print("powers adjoint vs FD rel %.2e" % ep, flush=True)
    assert ep < 1e-5


def test_adjoint_matches_oracle_differences(gpu):
    """Small case against central differences of the oracle (CPU FP64 path)."""
    rng = np.random.default_rng(21)
    top_r = rng.normal(scale=0.6, size=(5, 4)); bot_r = rng.normal(scale=0.6, size=(4, 4))
    top_p = rng.uniform(0.8, 2.0, 5); bot_p = rng.uniform(0.8, 2.0, 4)
    pts = rng.uniform(-1.2, 1.2, size=(300, 3))
    p = api.make_params(3, 2000.0, 0.9, 0.05, MAP_BLACK_HOLE)
    op = O.make_params(3, 2000.0, 0.9, 0.05, O.BLACK_HOLE)

    def J_cpu(tr, br, tp, bp):
        F = O.points_eval(pts, O.poly(tr, tp), O.poly(br, bp), op)
        return float(np.sum(np.where(np.isfinite(F), np.sin(F), 0.0)))

    r = api.Rational(api.POLYNOMIAL_4D(top_r, top_p), api.POLYNOMIAL_4D(bot_r, bot_p))
    d = torch.from_numpy(pts).cuda()
    F = api.points_eval(d, r, p, FP64)
    seed = torch.where(torch.isfinite(F), torch.cos(F), torch.zeros_like(F))
    gr, gp = api.points_adjoint(d, r, p, seed, FP64)
    gr = gr.cpu().numpy(); gp = gp.cpu().numpy()
    h = 1e-6
    fd = np.zeros_like(gr)
    for k, (arr, off) in enumerate(((top_r, 0), (bot_r, 5))):
        for i in range(arr.shape[0]):
            for c in range(4):
                a1 = arr.copy(); a1[i, c] += h
                a2 = arr.copy(); a2[i, c] -= h
                if k == 0:
                    fd[off + i, c] = (J_cpu(a1, bot_r, top_p, bot_p) - J_cpu(a2, bot_r, top_p, bot_p)) / (2 * h)
                else:
                    fd[off + i, c] = (J_cpu(top_r, a1, top_p, bot_p) - J_cpu(top_r, a2, top_p, bot_p)) / (2 * h)
    e = np.linalg.norm(gr - fd) / np.linalg.norm(fd)
    print("adjoint vs oracle FD rel %.2e" % e, flush=True)
    assert e < 1e-6
    fdp = []
    for i in range(5):
        t1 = top_p.copy(); t1[i] += h
        t2 = top_p.copy(); t2[i] -= h
        fdp.append((J_cpu(top_r, bot_r, t1, bot_p) - J_cpu(top_r, bot_r, t2, bot_p)) / (2 * h))
    assert np.allclose(gp[:5], fdp, rtol=1e-6, atol=1e-8 * np.max(np.abs(fdp)))


def test_autograd_function(gpu):
    w, pts = _c5(2000)
    p = api.make_params(3, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE)
    tr = torch.tensor(w.top.roots, requires_grad=True)
    tp = torch.tensor(w.top.powers, requires_grad=True)
    br = torch.tensor(w.bottom.roots, requires_grad=True)
    bp = torch.tensor(w.bottom.powers, requires_grad=True)
    F = G.julia_field(pts, tr, tp, br, bp, p)
    m = torch.isfinite(F)
    J = (F[m] ** 2).sum()
    J.backward()
    seed = torch.where(m, 2 * F, torch.zeros_like(F))
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    gr, gp = api.points_adjoint(pts, r, p, seed, FP64)
    assert torch.allclose(tr.grad, gr[:64].cpu()) and torch.allclose(bp.grad, gp[64:].cpu())
