"""GPU non-linear mesher (streamed slabs + bisection) against the oracle.

Triangles must be identical. Vertices come from FP64 bisection on re-evaluated
values; GPU (OCML) and CPU (glibc) values differ at the ulp level, so a
bisection step can in rare cases take the other half when a midpoint value is
within that difference of zero: at most 0.5% of vertices may differ, each by at
most one bisection interval (edge / 2^6); all others agree to 1e-12."""
import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE, MAP_LAGRANGE

pytestmark = pytest.mark.gpu


def _run(w, res, precision=FP64, chunk=64, iso=0.0):
    g = api.make_grid(res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    # a fresh workspace each time, so the buffers grow from empty across the chunks
    return api.nonlinear_mesh(g, r, p, iso, precision, 0.05, chunk, workspace=api.Workspace())


def _oracle(w, res, iso=0.0):
    return O.nonlinear_mc(O.make_grid(res, w.center, w.lengths), O.poly(w.top.roots, w.top.powers),
                          O.poly(w.bottom.roots, w.bottom.powers),
                          O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE), iso)


@pytest.mark.parametrize("cfg,res", [("c1", (48, 44, 40)), ("c2", (40, 40, 40))])
def test_nonlinear_mesher_vs_oracle(gpu, cfg, res):
    w = W.config1(8) if cfg == "c1" else W.config2(8)
    v, t = _run(w, res, FP64, chunk=9)
    ov, ot, bad = _oracle(w, res)
    assert bad == 0
    print("nlmc %s: %d verts %d tris" % (cfg, len(v), len(t)), flush=True)
    assert np.array_equal(t, ot)
    d = np.max(np.abs(v - ov), axis=1)
    edge = 3.0 / min(res)
    off = d > 1e-12
    assert off.mean() <= 0.005 and np.all(d <= edge / 2 ** 6 * 1.0001)


def test_chunking_and_precision_do_not_change_the_mesh(gpu):
    w = W.config2(8)
    res = (36, 33, 30)
    ref = _run(w, res, FP64, chunk=64)
    for prec, chunk in ((FP64, 1), (FP64, 5), (FP32_BAND, 7), (FP32_BAND, 64)):
        v, t = _run(w, res, prec, chunk)
        assert np.array_equal(t, ref[1]) and np.array_equal(v, ref[0]), (prec, chunk)


def test_slab_end_values_give_the_evaluated_vertices(gpu):
    """With FP32+band slabs and iso 0 the bisection takes its end values from
    the slab (their signs are the FP64 map's: cells with |F| < band hold the
    FP64 value itself); any other iso evaluates them. iso = 1e-300 changes no
    value above 1e-284 in magnitude, so both runs must give the same mesh bit
    for bit: the headline's 32+32 roots at 96^3 over three slabs."""
    w = W.headline(96)
    read = _run(w, w.res, FP32_BAND, chunk=32, iso=0.0)
    evaluated = _run(w, w.res, FP32_BAND, chunk=32, iso=1e-300)
    assert len(read[0]) > 50000
    assert np.array_equal(read[1], evaluated[1]) and np.array_equal(read[0], evaluated[0])


def test_device_output_is_the_host_mesh(gpu):
    """device_output=True returns the same mesh as the host arrays, bit for bit,
    including a mesh that spans several slabs (the slab-wide vertex lists)."""
    w = W.config2(8)
    res = (50, 47, 45)
    g = api.make_grid(res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    ws = api.Workspace()
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, 0.05, 8, workspace=ws)
    dv, dt = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, 0.05, 8, workspace=ws, device_output=True)
    assert isinstance(dv, torch.Tensor) and dv.is_cuda and dt.dtype == torch.int32
    assert len(v) > 1000
    assert np.array_equal(dv.cpu().numpy(), v) and np.array_equal(dt.cpu().numpy(), t)


def _setup(w, res, flags=MAP_BLACK_HOLE):
    g = api.make_grid(res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    return g, r, api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, flags)


def test_edge_cache_resume_rebuilds_the_same_mesh(gpu):
    w = W.config2(8)
    res = (40, 36, 32)
    g, r, p = _setup(w, res)
    ws = api.Workspace()
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP32_BAND, 0.05, 6, workspace=ws)
    edges, cells = api.nonlinear_mesh_cache(workspace=ws)
    assert len(edges) == len(v)
    # the cache is the reference's: edges straddle the sign change of (float)F in the
    # oracle's field, in vertex order, and the cells are every cell whose flag is not 0 or 255
    f = O.field_map(O.make_grid(res, w.center, w.lengths), O.poly(w.top.roots, w.top.powers),
                    O.poly(w.bottom.roots, w.bottom.powers),
                    O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
    pos = (f.astype(np.float32) > 0).reshape(res[2], res[1], res[0])
    x, y, z, ax = edges.T
    a = pos[z, y, x]
    b = pos[z + (ax == 2), y + (ax == 1), x + (ax == 0)]
    assert np.all(a != b)
    keys = ((z.astype(np.int64) * res[1] + y) * res[0] + x) * 3 + ax
    assert np.all(np.diff(keys) > 0)
    flag = np.zeros((res[2] - 1, res[1] - 1, res[0] - 1), dtype=np.int32)
    for k in range(8):
        dx, dy, dz = k >> 2, (k >> 1) & 1, k & 1
        flag += (1 << k) * pos[dz:res[2] - 1 + dz, dy:res[1] - 1 + dy, dx:res[0] - 1 + dx]
    zz, yy, xx = np.nonzero((flag != 0) & (flag != 255))
    assert np.array_equal(cells, np.c_[xx, yy, zz, flag[zz, yy, xx]].astype(np.int32))
    # rebuild from the cache in a fresh workspace: bit-identical
    v2, t2 = api.nonlinear_mesh_resume(g, r, p, edges, cells, workspace=api.Workspace())
    assert np.array_equal(v2, v) and np.array_equal(t2, t)
    # a cache missing an edge is rejected, not meshed
    with pytest.raises(RuntimeError):
        api.nonlinear_mesh_resume(g, r, p, np.delete(edges, len(edges) // 2, axis=0), cells,
                                  workspace=api.Workspace())
    # active cells with no edges at all: rejected on the host before any launch
    from quijibo_amd._lib import QJError
    with pytest.raises(QJError) as ei:
        api.nonlinear_mesh_resume(g, r, p, np.zeros((0, 4), np.int32), cells[:3], workspace=api.Workspace())
    assert ei.value.status == 1 and "no edges" in str(ei.value)


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_lagrange_mesher_vs_oracle(gpu, cfg):
    w = W.config1(8) if cfg == "c1" else W.config2(8)
    res = (36, 34, 32)
    g, r, p = _setup(w, res, MAP_BLACK_HOLE | MAP_LAGRANGE)
    v, t = api.nonlinear_mesh(g, r, p, 0.0, FP64, 0.05, 5, workspace=api.Workspace())
    og = O.make_grid(res, w.center, w.lengths)
    top, bot = O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers)
    op = O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE | O.LAGRANGE)
    ov, ot, bad = O.nonlinear_mc(og, top, bot, op)
    assert bad == 0 and np.array_equal(t, ot)
    d = np.max(np.abs(v - ov), axis=1)
    assert (d > 1e-12).mean() <= 0.005 and np.all(d <= 3.0 / min(res) / 2 ** 6 * 1.0001)
    # the field in the Lagrange order (FP64 only)
    F = api.field_eval(g, r, p, FP64).cpu().numpy()
    ref = O.field_map(og, top, bot, op)
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(F), fin)
    assert np.max(np.abs(F[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))) <= 1e-9
    if cfg == "c2":
        std = O.field_map(og, top, bot, O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
        assert np.max(np.abs(std[fin] - ref[fin])) > 1e-3  # a different map from the factored order
    with pytest.raises(RuntimeError):
        api.field_eval(g, r, p, FP32_BAND)


def test_iterate_on_zero_and_forced_black_hole(gpu):
    """An iterate that underflows to exactly 0 at a bisection midpoint: the
    Lagrange mesher's value is FLT_MAX, the factored mesher's -inf, as in the
    oracle; and the mesher applies the black-hole stop whatever the flags say."""
    from test_oracle_mc import edge_vertex, zero_iterate_case
    og, p0, p1, m, top, bot, c = zero_iterate_case()
    g = api.make_grid((6, 6, 6), (0.0, 0.0, 0.0), (0.6, 0.6, 0.6))
    r = api.Rational(api.POLYNOMIAL_4D(top, [3.0]), api.POLYNOMIAL_4D(bot, [1.0]))
    for flags, oflags in ((MAP_BLACK_HOLE | MAP_LAGRANGE, O.BLACK_HOLE | O.LAGRANGE), (MAP_BLACK_HOLE, O.BLACK_HOLE),
                          (MAP_LAGRANGE, O.BLACK_HOLE | O.LAGRANGE), (0, O.BLACK_HOLE)):
        v, t = api.nonlinear_mesh(g, r, api.make_params(1, 2000.0, c, 0.0, flags), 0.0, FP64, 0.05, 2,
                                  workspace=api.Workspace())
        ov, ot, bad = O.nonlinear_mc(og, O.poly(top, [3.0]), O.poly(bot, [1.0]),
                                     O.make_params(1, 2000.0, c, 0.0, oflags))
        assert bad == 0 and np.array_equal(t, ot), flags
        assert np.max(np.abs(v - ov)) <= 1e-12, flags
        x = edge_vertex(v, p0, p1)[0]
        if flags & MAP_LAGRANGE:
            assert p0[0] < x <= m[0]
        else:
            assert m[0] <= x < p1[0]


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_lagrange_points_eval_matches_oracle(gpu, cfg):
    """qj_points_eval in FP64 with the Lagrange order (the polar-form point kernel)."""
    w = W.config1(8) if cfg == "c1" else W.config2(8)
    rng = np.random.default_rng(17)
    pts = rng.uniform(-1.4, 1.4, size=(5000, 3))
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers * 1.1),
                     api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers * 1.1))
    p = api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, MAP_BLACK_HOLE | MAP_LAGRANGE)
    F = api.points_eval(torch.from_numpy(pts).cuda(), r, p, FP64).cpu().numpy()
    ref = O.points_eval(pts, O.poly(w.top.roots, w.top.powers * 1.1), O.poly(w.bottom.roots, w.bottom.powers * 1.1),
                        O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE | O.LAGRANGE))
    fin = np.isfinite(ref)
    assert np.array_equal(np.isfinite(F), fin)
    assert np.max(np.abs(F[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))) <= 1e-9
    if cfg == "c2":
        std = O.points_eval(pts, O.poly(w.top.roots, w.top.powers * 1.1),
                            O.poly(w.bottom.roots, w.bottom.powers * 1.1),
                            O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
        assert np.max(np.abs(std[fin] - ref[fin])) > 1e-3


@pytest.mark.parametrize("cfg", ["c1", "c2"])
def test_nonlinear_mesher_on_a_field(gpu, cfg):
    """TRIANGLE_MESH(field, top, bottom, c, maxIter, slice, iso): cases from the
    given field, vertices by bisection on the map; against the oracle's
    restatement of the same constructor on the same field."""
    w = W.config1(8) if cfg == "c1" else W.config2(8)
    res = (38, 35, 33)
    g, r, p = _setup(w, res)
    field = api.field_eval(g, r, p, FP64)
    # a field of its own: the map's field with a few signs flipped in one corner,
    # so the cases are the field's and not the map's
    f = field.clone()
    f.view(res[2], res[1], res[0])[:6, :6, :6] *= -1.0
    v, t = api.nonlinear_mesh_field(f, g, r, p, 0.0, workspace=api.Workspace(), allow_inconsistent=True)
    og = O.make_grid(res, w.center, w.lengths)
    ov, ot, bad = O.nonlinear_mc_field(og, f.cpu().numpy(), O.poly(w.top.roots, w.top.powers),
                                       O.poly(w.bottom.roots, w.bottom.powers),
                                       O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
    print("field mesher %s: %d vertices %d triangles, %d edges not straddling" % (cfg, len(v), len(t), bad),
          flush=True)
    assert np.array_equal(t, ot) and len(t) > 100
    d = np.max(np.abs(v - ov), axis=1)
    assert (d > 1e-12).mean() <= 0.005 and np.all(d <= 3.0 / min(res) / 2 ** 6 * 1.0001)
    # on the map's own field the constructor gives the slab mesher's mesh
    v2, t2 = api.nonlinear_mesh_field(field, g, r, p, 0.0, workspace=api.Workspace())
    v3, t3 = api.nonlinear_mesh(g, r, p, 0.0, FP64, 0.05, 64, workspace=api.Workspace())
    assert np.array_equal(t2, t3) and np.array_equal(v2, v3)
