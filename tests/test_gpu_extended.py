"""QJ_MAP_EXTENDED_RANGE (the reference's default-build range, Real = long
double) on the GPU, against the oracle's long double map and the complex
plane's log-space closed form.

Tolerances: FP64 precisions |dF| <= 1e-9 max(1, |F|) and the same
non-finite cells; FP32 + FP64 band: zero sign flips of (float)F > 0, the same
non-finite cells, FP64 tolerance inside the band; meshes: identical
triangles (linear mesher bit for bit on the same field)."""
import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api, server as S, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, FP64_EXACT, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE

pytestmark = pytest.mark.gpu

EXT = MAP_BLACK_HOLE | MAP_EXTENDED_RANGE


def _rational(w):
    return api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers, w.top.power_scalar),
                        api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar))


def _gpu(w, prec, flags=EXT, band=0.05, res=None, center=None, lengths=None, ws=None):
    g = api.make_grid(res or w.res, center or w.center, lengths or w.lengths)
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags)
    return api.field_eval(g, _rational(w), p, prec, band, workspace=ws).cpu().numpy()


def _oracle(w, flags=O.BLACK_HOLE | O.EXTENDED, res=None, center=None, lengths=None):
    g = O.make_grid(res or w.res, center or w.center, lengths or w.lengths)
    return O.field_map(g, O.poly(w.top.roots, w.top.powers, w.top.power_scalar),
                       O.poly(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar),
                       O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags))


def _close64(a, ref, rtol=1e-9):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(a), fin) and np.array_equal(np.isnan(a), np.isnan(ref))
    e = np.abs(a[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))
    print("ext fp64: max rel %.2e over %d cells, %d non-finite" % (e.max(), fin.sum(), (~fin).sum()), flush=True)
    assert e.max() <= rtol


def _flips(a, b):
    return int(np.count_nonzero((a.astype(np.float32) > 0) != (b.astype(np.float32) > 0)))


def test_plane_closed_form_all_precisions(gpu):
    from test_oracle_extended import plane_log_space, plane_workload
    w = plane_workload()
    ref = plane_log_space(w)
    assert (~np.isfinite(_gpu(w, FP64, MAP_BLACK_HOLE))).sum() > 50  # FP64 range overflows here
    for prec in (FP64, FP64_EXACT):
        _close64(_gpu(w, prec), ref)
    f32 = _gpu(w, FP32_BAND)
    assert np.isfinite(f32).all() and _flips(f32, ref) == 0
    inband = np.abs(ref) < 0.04
    assert np.all(np.abs(f32[inband] - ref[inband]) <= 1e-9 * np.maximum(1.0, np.abs(ref[inband])))


def test_torus_roots_against_the_long_double_oracle(gpu):
    """128 + 128 torus roots (the 1024^3 workload's): FP64 range gives NaN and
    -inf cells; the extended range gives the long double values."""
    w = W.config3(1024)
    c, L = W.tile_grid(w, 2, 1, 1, 256)
    res = (40, 36, 6)
    zc = (c[0], c[1], c[2] + (120 + 3 - 128) * L[2] / 256)
    zl = (L[0], L[1], 6 * L[2] / 256)
    ref = _oracle(w, res=res, center=zc, lengths=zl)
    plain = _oracle(w, flags=O.BLACK_HOLE, res=res, center=zc, lengths=zl)
    print("torus window: fp64 oracle %d non-finite, long double %d" % ((~np.isfinite(plain)).sum(),
                                                                        (~np.isfinite(ref)).sum()), flush=True)
    for prec in (FP64, FP64_EXACT):
        _close64(_gpu(w, prec, res=res, center=zc, lengths=zl), ref)
    f32 = _gpu(w, FP32_BAND, res=res, center=zc, lengths=zl)
    assert _flips(f32, ref) == 0 and np.array_equal(np.isfinite(f32), np.isfinite(ref))


def test_points_and_the_nonlinear_mesher(gpu):
    from test_oracle_extended import plane_workload
    w = plane_workload(res=(28, 26, 24))
    w.top.roots[:, 2] = np.linspace(-0.3, 0.3, w.top.n)  # off the plane: general quaternions
    w.lengths = (6.0, 5.4, 4.8)
    rng = np.random.default_rng(3)
    pts = rng.uniform(-2.5, 2.5, size=(3000, 3))
    r = _rational(w)
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, EXT)
    got = api.points_eval(torch.from_numpy(pts).cuda(), r, p, FP64).cpu().numpy()
    ref = O.points_eval(pts, O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers),
                        O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, O.BLACK_HOLE | O.EXTENDED))
    _close64(got, ref)
    g = api.make_grid(w.res, w.center, w.lengths)
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, 0.05, 5, workspace=api.Workspace())
    ov, ot, bad = O.nonlinear_mc(O.make_grid(w.res, w.center, w.lengths), O.poly(w.top.roots, w.top.powers),
                                 O.poly(w.bottom.roots, w.bottom.powers),
                                 O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice,
                                               O.BLACK_HOLE | O.EXTENDED))
    print("ext mesher: %d vertices %d triangles" % (len(v), len(t)), flush=True)
    assert bad == 0 and len(t) > 100 and np.array_equal(t, ot)
    d = np.max(np.abs(v - ov), axis=1)
    assert (d > 1e-12).mean() <= 0.005 and np.all(d <= max(w.lengths) / min(w.res) / 2 ** 6 * 1.0001)
    assert np.isfinite(v).all()


def test_request_loop_and_cross_check(gpu):
    """A request with the flag through the request loop equals qj_field_eval;
    the sampled cross-check compares against the extended FP64 method."""
    w = W.config3(48)
    srv = S.Server(hbm_budget_bytes=64 << 20, chunk_voxels=4096, chunks_per_launch=16, precision=FP32_BAND)
    rid = srv.submit(w, keep_field=True, flags=EXT)
    c = srv.wait_all(1, timeout=60)
    assert len(c) == 1 and c[0].status == 0
    got = srv.fetch(c[0], device=False)
    srv.release(rid)
    srv.close()
    ws = api.Workspace()
    ws.check(4096)
    direct = _gpu(w, FP32_BAND, ws=ws)
    assert np.array_equal(got, direct, equal_nan=True)
    st = ws.stats()
    print("ext cross-check:", {k: st[k] for k in st if k.startswith("check")}, flush=True)
    assert st["checked"] == 4096 and st["check_sign_flips"] == 0 and st["check_class_mismatch"] == 0
    assert np.isfinite(direct).all()
    plain = _gpu(w, FP32_BAND, flags=MAP_BLACK_HOLE)
    assert not np.isfinite(plain).all()  # the same roots in FP64 range


def test_clamp_then_mesh_has_no_nan_vertices(gpu):
    """The torus window of the clamp-then-mesh test in the extended range: no
    cell needs clamping, and the mesh equals the oracle's on the same field."""
    w = W.config3()
    c, L = W.tile_grid(w, 2, 2, 2, 128)
    res = (128, 128, 128)
    g = api.make_grid(res, c, L)
    p = api.make_params(w.max_iter, 2000.0, 1.0, 0.0, EXT)
    raw = api.field_eval(g, _rational(w), p, FP32_BAND, 0.05)
    assert torch.isfinite(raw).all()
    F = api.FIELD_3D(*res, center=c, lengths=L, data=raw.clone())
    F.clampInfsToMean()
    F.setBorders(F.totalCells())
    v, t = api.marching_cubes(F.data.view(-1), F.grid())
    ov, ot = O.marching_cubes(O.make_grid(res, c, L), F.numpy())
    assert np.array_equal(v, ov) and np.array_equal(t, ot) and np.isfinite(v).all() and len(t) > 1000
