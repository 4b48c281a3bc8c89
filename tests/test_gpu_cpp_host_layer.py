"""The C++ host layer (POLYNOMIAL_4D, FIELD_3D, TRIANGLE_MESH, OPTIMIZE_3D over
the C ABI) driven from a C++ program: inputs written as a .o3d file by the
Python format writer, outputs read back from the FIELD_3D and OBJ files the
C++ classes wrote, compared with the oracle."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from quijibo_amd import formats as FM, workloads as W

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "host_layer_driver")


def _sphere_distance(res, center, lengths, r=0.8):
    ax = [center[i] - lengths[i] / 2 + (np.arange(res[i]) + 0.5) * lengths[i] / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - r).reshape(-1)


def _o3d(tmp_path, w, res):
    """The workload as a .o3d; its distance field (a sphere) sets the grid that
    OPTIMIZE_3D::read gives the fractal (setScoreFields, OPTIMIZE_3D.cpp:3442-3471)."""
    s = FM.O3D(res, w.center, w.lengths, FM.Polynomial4D(w.top.roots, w.top.powers, w.top.power_scalar),
               FM.Polynomial4D(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar), w.exp_scaling, 0,
               FM.Field3D(res, w.center, w.lengths),
               FM.Field3D(res, w.center, w.lengths, _sphere_distance(res, w.center, w.lengths)))
    p = tmp_path / "in.o3d"
    FM.write_o3d(p, s)
    return p


def _run(args):
    print("driver:", " ".join(args[1:]), flush=True)
    r = subprocess.run([DRIVER] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def _oracle_map(w, res):
    g = O.make_grid(res, w.center, w.lengths)
    t, b = O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers)
    p = O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE)
    return g, t, b, p


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_map_and_meshes(gpu, tmp_path, cfg):
    w = W.config1(8) if cfg == "c1" else W.config2(8)
    res = (40, 36, 32)
    _run(["map", str(_o3d(tmp_path, w, res)), str(w.max_iter), "0", str(tmp_path)])
    g, t, b, p = _oracle_map(w, res)
    field = FM.read_field(tmp_path / "field.bin")
    assert field.res == res and field.center == tuple(w.center)
    ref = O.field_map(g, t, b, p)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(field.data), fin)
    err = np.abs(field.data[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))
    assert err.max() <= 1e-9
    # linear mesher on the field the C++ layer produced: bit-exact vs the oracle's MC
    lv, lt = FM.read_obj(tmp_path / "linear.obj")
    ov, ot = O.marching_cubes(g, field.data)
    assert np.array_equal(lt, ot) and np.max(np.abs(lv - ov)) <= 1e-15
    # clampInfsToMean + setBorders(totalCells) through the C++ FIELD_3D, then its mesh
    cf = FM.read_field(tmp_path / "clamped.bin")
    rf, _ = O.clamp_infs_to_mean(field.data)
    rf = O.set_borders(g, rf, float(np.prod(res)))
    same = np.isfinite(field.data)
    assert np.array_equal(cf.data[same], rf[same])
    assert np.allclose(cf.data[~same], rf[~same], rtol=1e-12, atol=0, equal_nan=True)
    cv, ct = FM.read_obj(tmp_path / "clamped.obj")
    ov, ot = O.marching_cubes(g, cf.data)
    assert np.array_equal(ct, ot) and np.max(np.abs(cv - ov), initial=0.0) <= 1e-15
    # non-linear mesher (slabs of 7 slices): same triangles as the oracle
    nv, nt = FM.read_obj(tmp_path / "nonlinear.obj")
    pv, pt, bad = O.nonlinear_mc(g, t, b, p)
    assert bad == 0 and np.array_equal(nt, pt)
    d = np.max(np.abs(nv - pv), axis=1)
    assert (d > 1e-15).mean() <= 0.005 and np.all(d <= 3.0 / min(res) / 2 ** 6 * 1.0001)
    # the edge cache the constructor wrote, and the mesh rebuilt from it
    edges, cells = FM.read_edge_cache(tmp_path / "mesh.cache")
    assert len(edges) == len(nv) and len(cells) > 0
    assert (tmp_path / "resumed.obj").read_bytes() == (tmp_path / "nonlinear.obj").read_bytes()
    # the Lagrange constructor
    lv, lt = FM.read_obj(tmp_path / "lagrange.obj")
    lp = O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE | O.LAGRANGE)
    qv, qt, bad = O.nonlinear_mc(g, t, b, lp)
    assert bad == 0 and np.array_equal(lt, qt)
    d = np.max(np.abs(lv - qv), axis=1)
    assert (d > 1e-15).mean() <= 0.005 and np.all(d <= 3.0 / min(res) / 2 ** 6 * 1.0001)


def test_fp32_band_map_keeps_the_topology(gpu, tmp_path):
    w = W.config2(8)
    res = (40, 36, 32)
    (tmp_path / "a").mkdir()
    (tmp_path / "b").mkdir()
    p = _o3d(tmp_path, w, res)
    _run(["map", str(p), str(w.max_iter), "0", str(tmp_path / "a")])
    _run(["map", str(p), str(w.max_iter), "1", str(tmp_path / "b")])
    _, t64 = FM.read_obj(tmp_path / "a" / "linear.obj")
    _, t32 = FM.read_obj(tmp_path / "b" / "linear.obj")
    assert np.array_equal(t64, t32)
    v64, n64 = FM.read_obj(tmp_path / "a" / "nonlinear.obj")
    v32, n32 = FM.read_obj(tmp_path / "b" / "nonlinear.obj")
    assert np.array_equal(n64, n32) and np.array_equal(v64, v32)


def test_ap_score_and_gradients(gpu, tmp_path):
    w = W.config2(8)
    res = (24, 22, 20)
    ax = [-1.5 + (np.arange(res[i]) + 0.5) * 3.0 / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    weight = (1.0 / (np.abs(np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.8) + 0.05)).reshape(-1)
    FM.write_field(tmp_path / "w.field", FM.Field3D(res, w.center, w.lengths, weight))
    out = json.loads(_run(["ap", str(_o3d(tmp_path, w, res)), str(tmp_path / "w.field"), "10.0", "0.0"]))
    g = O.make_grid(res, w.center, w.lengths)
    cpu = O.ap_score_gradients(g, O.poly(w.top.roots, w.top.powers, w.top.power_scalar),
                               O.poly(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar),
                               O.make_params(1, 20.0, w.exp_scaling, w.slice, O.SKIP_ORIGIN | O.AP_REPASS), g,
                               weight)
    for k in ("score", "grad_exp", "grad_top_scalar", "grad_bottom_scalar", "grad_top", "grad_bottom"):
        a = np.atleast_1d(np.asarray(out[k], dtype=np.float64))
        b = np.atleast_1d(np.asarray(cpu[k], dtype=np.float64))
        assert np.max(np.abs(a - b)) <= 1e-9 * np.max(np.abs(b)), k
    assert out["adaptive_cells"] == cpu["overflowed"]


def test_blmvm_exponent_fit_follows_the_oracle(gpu, tmp_path):
    """The shape optimizer's exponent fit (FormScaledFunctionGradient packing,
    toothOptimizer.cpp:65-110) with the GPU AP pass as the callback, against
    the oracle solver driven by the oracle's AP score and gradients."""
    from oracle import blmvm as B
    w = W.config2(8)
    rng = np.random.default_rng(7)
    w.top.roots = w.top.roots[:6]
    w.top.powers = rng.choice([1.0, 2.0], 6)
    w.bottom.roots = w.bottom.roots[:5]
    w.bottom.powers = rng.choice([1.0, 2.0], 5)
    res = (20, 18, 16)
    ax = [-1.5 + (np.arange(res[i]) + 0.5) * 3.0 / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    weight = (np.sqrt(X ** 2 + Y ** 2 + Z ** 2) - 0.9).reshape(-1)
    weight = np.where(weight < 0, -1.0, 1.0) / (np.abs(weight) + 0.1)
    weight = weight / np.abs(weight).sum()
    FM.write_field(tmp_path / "w.field", FM.Field3D(res, w.center, w.lengths, weight))
    out = json.loads(_run(["optimize", str(_o3d(tmp_path, w, res)), str(tmp_path / "w.field"), "10.0", "6"]))
    g = O.make_grid(res, w.center, w.lengths)
    nt, nb = len(w.top.powers), len(w.bottom.powers)

    def fg(state):
        c = float(np.exp(state[0]))
        bp = np.array(state[1:1 + nb])
        tp = np.array(state[1 + nb:1 + nb + nt])
        r = O.ap_score_gradients(g, O.poly(w.top.roots, tp, state[-2]), O.poly(w.bottom.roots, bp, state[-1]),
                                 O.make_params(1, 20.0, c, w.slice, O.SKIP_ORIGIN | O.AP_REPASS), g, weight, 10.0, 0.0)
        grad = [r["grad_exp"]] + list(r["grad_bottom"]) + list(r["grad_top"]) + [r["grad_top_scalar"],
                                                                                 r["grad_bottom_scalar"]]
        return r["score"], [float(v) for v in grad]

    x0 = [0.0] + list(w.bottom.powers) + list(w.top.powers) + [1.0, 1.0]
    n = len(x0)
    lo = [-3.4028234663852886e38] + [0.0] * (n - 1)
    hi = [3.4028234663852886e38] * n
    x, hist, it, reason = B.optimize_blmvm(x0, hi, lo, fg, max_it=6)
    print("gpu history", out["history"], "\noracle history", hist, flush=True)
    assert out["iterations"] == it and out["reason"] == reason
    assert np.allclose(out["history"], hist, rtol=1e-7, atol=1e-12)
    # GPU and CPU gradients agree to ~1e-12; six quasi-Newton steps on a flat
    # objective amplify that to ~1e-6 in the state (the objective agrees to ~1e-8)
    assert np.allclose(out["x"], x, rtol=2e-5, atol=1e-8)
    assert hist[-1] < hist[0]


def test_read_sets_up_the_score_fields(gpu, tmp_path):
    """OPTIMIZE_3D::read's set-up (OPTIMIZE_3D.cpp:3348-3388): the grid from the
    distance field, the fractal zeroed on it, the curvature field a copy of
    the distance field, the weight field invertedDistanceNormalized(distance)
    and its absolute sum, against the oracle."""
    w = W.config2(8)
    res = (22, 20, 18)
    center, lengths = (0.2, -0.1, 0.05), (2.6, 2.4, 2.2)
    dist = _sphere_distance(res, center, lengths, 0.7)
    s = FM.O3D((5, 5, 5), (0.0, 0.0, 0.0), (3.0, 3.0, 3.0),
               FM.Polynomial4D(w.top.roots, w.top.powers), FM.Polynomial4D(w.bottom.roots, w.bottom.powers), 1.0, 0,
               FM.Field3D((5, 5, 5), (0.0, 0.0, 0.0), (3.0, 3.0, 3.0), np.arange(125.0)),
               FM.Field3D(res, center, lengths, dist), FM.Field3D((2, 2, 2), data=np.ones(8)))
    FM.write_o3d(tmp_path / "in.o3d", s)
    out = json.loads(_run(["readsetup", str(tmp_path / "in.o3d"), str(tmp_path)]))
    assert tuple(out["res"]) == res
    fr = FM.read_field(tmp_path / "fractal.bin")
    assert fr.res == res and fr.center == center and fr.lengths == lengths and not fr.data.any()
    cv = FM.read_field(tmp_path / "curvature.bin")
    assert cv.res == res and np.array_equal(cv.data, dist)
    ref, a = O.inverted_distance(O.make_grid(res, center, lengths), dist, normalized=True)
    wt = FM.read_field(tmp_path / "weight.bin")
    inv = FM.read_field(tmp_path / "inverse.bin")
    assert np.array_equal(wt.data, inv.data)
    # the layer numbers are exact; the normalising sums are double-double on
    # the device and serial in the oracle, so the scale factors differ by up
    # to n u (u = 2^-53) relative
    raw, _ = O.inverted_distance(O.make_grid(res, center, lengths), dist)
    assert np.array_equal(np.sign(wt.data), np.sign(raw))
    tol = ref.size * 2.0 ** -52
    assert np.max(np.abs(wt.data - ref) / np.abs(ref)) <= tol
    assert abs(out["inverse_sum"] - a) <= tol
