"""Oracle (test infrastructure only): ctypes bindings to the C restatement in qo.c.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this
package, as the reference results the GPU path is compared with and as the CPU
baseline. The shipped library never uses it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

BLACK_HOLE = 1
SKIP_ORIGIN = 2
AP_REPASS = 4
LAGRANGE = 8
EXTENDED = 16


class Grid(ctypes.Structure):
    _fields_ = [("res", ctypes.c_int32 * 3), ("center", ctypes.c_double * 3),
                ("lengths", ctypes.c_double * 3), ("rotation", ctypes.c_double * 4)]


class Poly(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("roots", ctypes.POINTER(ctypes.c_double)),
                ("powers", ctypes.POINTER(ctypes.c_double)), ("power_scalar", ctypes.c_double)]


class Params(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("escape", ctypes.c_double),
                ("exp_scaling", ctypes.c_double), ("slice", ctypes.c_double),
                ("flags", ctypes.c_uint32)]


class APResult(ctypes.Structure):
    _fields_ = [("score", ctypes.c_double), ("grad_exp", ctypes.c_double),
                ("grad_top_scalar", ctypes.c_double), ("grad_bottom_scalar", ctypes.c_double),
                ("overflowed", ctypes.c_int64)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.qo_qmul.argtypes = [dp, dp, dp]
        L.qo_qdiv.argtypes = [dp, dp, dp]
        L.qo_qpow.argtypes = [dp, ctypes.c_double, dp]
        L.qo_qlog.argtypes = [dp, dp]
        L.qo_qmag.argtypes = [dp]
        L.qo_qmag.restype = ctypes.c_double
        L.qo_poly_eval.argtypes = [ctypes.POINTER(Poly), dp, dp]
        L.qo_cell_center.argtypes = [ctypes.POINTER(Grid), ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
        L.qo_field_map.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Poly), ctypes.POINTER(Poly),
                                   ctypes.POINTER(Params), dp, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
        L.qo_field_map_zrange.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Poly), ctypes.POINTER(Poly),
                                          ctypes.POINTER(Params), ctypes.c_int, ctypes.c_int, dp, ctypes.c_void_p,
                                          ctypes.c_void_p, ctypes.c_int]
        L.qo_points_eval.argtypes = [dp, ctypes.c_int64, ctypes.POINTER(Poly), ctypes.POINTER(Poly),
                                     ctypes.POINTER(Params), dp, ctypes.c_int]
        L.qo_marching_cubes.argtypes = [ctypes.POINTER(Grid), dp, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p]
        L.qo_marching_cubes.restype = ctypes.c_int
        L.qo_nonlinear_mc.argtypes = [ctypes.POINTER(Grid), ctypes.POINTER(Poly), ctypes.POINTER(Poly),
                                      ctypes.POINTER(Params), ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
                                      ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p, ctypes.c_void_p]
        L.qo_nonlinear_mc.restype = ctypes.c_int64
        L.qo_nonlinear_mc_field.argtypes = [ctypes.POINTER(Grid), ctypes.c_void_p, ctypes.POINTER(Poly),
                                            ctypes.POINTER(Poly), ctypes.POINTER(Params), ctypes.c_double,
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                            ctypes.c_void_p, ctypes.c_void_p]
        L.qo_nonlinear_mc_field.restype = ctypes.c_int64
        L.qo_field_lookup.argtypes = [ctypes.POINTER(Grid), dp, dp]
        L.qo_field_lookup.restype = ctypes.c_double
        L.qo_ap_score_gradients.argtypes = [
            ctypes.POINTER(Grid), ctypes.POINTER(Poly), ctypes.POINTER(Poly), ctypes.POINTER(Params),
            ctypes.POINTER(Grid), dp, ctypes.c_double, ctypes.c_double, dp, dp, dp,
            ctypes.POINTER(APResult), ctypes.c_int]
        L.qo_clamp_infs_to_mean.argtypes = [dp, ctypes.c_int64]
        L.qo_clamp_infs_to_mean.restype = ctypes.c_double
        L.qo_clamp_nans_to_mean.argtypes = [dp, ctypes.c_int64]
        L.qo_clamp_nans_to_mean.restype = ctypes.c_double
        L.qo_set_borders.argtypes = [ctypes.POINTER(Grid), dp, ctypes.c_double]
        L.qo_inverted_distance.argtypes = [ctypes.POINTER(Grid), dp, dp]
        L.qo_inverted_distance.restype = ctypes.c_int
        L.qo_inverted_distance_normalized.argtypes = [ctypes.POINTER(Grid), dp, dp, dp]
        L.qo_inverted_distance_normalized.restype = ctypes.c_int
        L.qo_weight_score.argtypes = [ctypes.POINTER(Grid), dp, ctypes.POINTER(Grid), dp, ctypes.c_double,
                                      ctypes.c_double, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_grid(res, center=(0.0, 0.0, 0.0), lengths=(3.0, 3.0, 3.0), rotation=(1.0, 0.0, 0.0, 0.0)):
    if isinstance(res, int):
        res = (res, res, res)
    g = Grid()
    g.res[:] = list(res)
    g.center[:] = list(center)
    g.lengths[:] = list(lengths)
    g.rotation[:] = list(rotation)
    return g


class _PolyHolder:
    """Keeps the numpy arrays alive for as long as the ctypes struct is used."""

    def __init__(self, roots, powers=None, power_scalar=1.0):
        self.roots = np.ascontiguousarray(np.asarray(roots, dtype=np.float64).reshape(-1, 4))
        n = self.roots.shape[0]
        self.powers = np.ascontiguousarray(np.ones(n) if powers is None else np.asarray(powers, dtype=np.float64))
        self.p = Poly(n, _dp(self.roots), _dp(self.powers), float(power_scalar))


def poly(roots, powers=None, power_scalar=1.0):
    return _PolyHolder(roots, powers, power_scalar)


def make_params(max_iter=3, escape=2000.0, exp_scaling=1.0, slice_=0.0, flags=BLACK_HOLE):
    return Params(int(max_iter), float(escape), float(exp_scaling), float(slice_), int(flags))


def field_map(grid, top, bottom, params, nthreads=0, want_iters=False, want_overflow=False, init=None):
    n = grid.res[0] * grid.res[1] * grid.res[2]
    out = np.zeros(n) if init is None else np.array(init, dtype=np.float64).reshape(-1).copy()
    iters = np.zeros(n, dtype=np.int32) if want_iters else None
    ovf = np.zeros(n, dtype=np.uint8) if want_overflow else None
    lib().qo_field_map(ctypes.byref(grid), ctypes.byref(top.p), ctypes.byref(bottom.p), ctypes.byref(params),
                       _dp(out), iters.ctypes.data if iters is not None else None,
                       ovf.ctypes.data if ovf is not None else None, int(nthreads))
    res = [out]
    if want_iters:
        res.append(iters)
    if want_overflow:
        res.append(ovf)
    return res[0] if len(res) == 1 else tuple(res)


def field_map_zrange(grid, top, bottom, params, z0, z1, nthreads=0):
    n = grid.res[0] * grid.res[1] * (z1 - z0)
    out = np.zeros(n)
    lib().qo_field_map_zrange(ctypes.byref(grid), ctypes.byref(top.p), ctypes.byref(bottom.p), ctypes.byref(params),
                              int(z0), int(z1), _dp(out), None, None, int(nthreads))
    return out


def points_eval(points, top, bottom, params, nthreads=0):
    pts = np.ascontiguousarray(np.asarray(points, dtype=np.float64).reshape(-1, 3))
    out = np.empty(pts.shape[0])
    lib().qo_points_eval(_dp(pts), pts.shape[0], ctypes.byref(top.p), ctypes.byref(bottom.p),
                         ctypes.byref(params), _dp(out), int(nthreads))
    return out


def marching_cubes(grid, field):
    f = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(-1))
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = lib()
    if L.qo_marching_cubes(ctypes.byref(grid), _dp(f), ctypes.byref(nv), ctypes.byref(nt), None, None):
        raise MemoryError("oracle marching cubes")
    verts = np.empty((nv.value, 3))
    tris = np.empty((nt.value, 3), dtype=np.int32)
    if L.qo_marching_cubes(ctypes.byref(grid), _dp(f), ctypes.byref(nv), ctypes.byref(nt),
                           verts.ctypes.data, tris.ctypes.data):
        raise MemoryError("oracle marching cubes")
    return verts, tris[: nt.value].copy()


def nonlinear_mc(grid, top, bottom, params, iso=0.0):
    """Returns (vertices, triangles, n_bad_edges)."""
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = lib()
    L.qo_nonlinear_mc(ctypes.byref(grid), ctypes.byref(top.p), ctypes.byref(bottom.p), ctypes.byref(params),
                      float(iso), ctypes.byref(nv), ctypes.byref(nt), None, None)
    verts = np.empty((nv.value, 3))
    tris = np.empty((max(nt.value, 1), 3), dtype=np.int32)
    bad = L.qo_nonlinear_mc(ctypes.byref(grid), ctypes.byref(top.p), ctypes.byref(bottom.p), ctypes.byref(params),
                            float(iso), ctypes.byref(nv), ctypes.byref(nt), verts.ctypes.data, tris.ctypes.data)
    return verts, tris[: nt.value].copy(), int(bad)


def nonlinear_mc_field(grid, field, top, bottom, params, iso=0.0):
    """The non-linear mesher on a given field; returns (vertices, triangles, n_bad_edges)."""
    f = np.ascontiguousarray(np.asarray(field, dtype=np.float64).reshape(-1))
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = lib()
    L.qo_nonlinear_mc_field(ctypes.byref(grid), f.ctypes.data, ctypes.byref(top.p), ctypes.byref(bottom.p),
                            ctypes.byref(params), float(iso), ctypes.byref(nv), ctypes.byref(nt), None, None)
    verts = np.empty((nv.value, 3))
    tris = np.empty((max(nt.value, 1), 3), dtype=np.int32)
    bad = L.qo_nonlinear_mc_field(ctypes.byref(grid), f.ctypes.data, ctypes.byref(top.p), ctypes.byref(bottom.p),
                                  ctypes.byref(params), float(iso), ctypes.byref(nv), ctypes.byref(nt),
                                  verts.ctypes.data, tris.ctypes.data)
    return verts, tris[: nt.value].copy(), int(bad)


def cell_center(grid, x, y, z):
    out = np.empty(3)
    lib().qo_cell_center(ctypes.byref(grid), x, y, z, _dp(out))
    return out


def field_lookup(grid, data, pos):
    d = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1))
    p = np.ascontiguousarray(np.asarray(pos, dtype=np.float64))
    return lib().qo_field_lookup(ctypes.byref(grid), _dp(d), _dp(p))


def ap_score_gradients(grid, top, bottom, params, wgrid, weight, tanh_alpha=10.0, tanh_threshold=0.0,
                       nthreads=0, field_init=None):
    n = grid.res[0] * grid.res[1] * grid.res[2]
    field = np.zeros(n) if field_init is None else np.array(field_init, dtype=np.float64).reshape(-1).copy()
    w = np.ascontiguousarray(np.asarray(weight, dtype=np.float64).reshape(-1))
    gt = np.zeros(max(top.p.n, 1))
    gb = np.zeros(max(bottom.p.n, 1))
    r = APResult()
    lib().qo_ap_score_gradients(ctypes.byref(grid), ctypes.byref(top.p), ctypes.byref(bottom.p),
                                ctypes.byref(params), ctypes.byref(wgrid), _dp(w), float(tanh_alpha),
                                float(tanh_threshold), _dp(field), _dp(gt), _dp(gb), ctypes.byref(r),
                                int(nthreads))
    return dict(score=r.score, grad_exp=r.grad_exp, grad_top_scalar=r.grad_top_scalar,
                grad_bottom_scalar=r.grad_bottom_scalar, overflowed=r.overflowed,
                grad_top=gt[: top.p.n].copy(), grad_bottom=gb[: bottom.p.n].copy(), field=field)


def qmul(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64); b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(4); lib().qo_qmul(_dp(a), _dp(b), _dp(out)); return out


def qdiv(a, b):
    a = np.ascontiguousarray(a, dtype=np.float64); b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(4); lib().qo_qdiv(_dp(a), _dp(b), _dp(out)); return out


def qpow(a, e):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(4); lib().qo_qpow(_dp(a), float(e), _dp(out)); return out


def qlog(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(4); lib().qo_qlog(_dp(a), _dp(out)); return out


def qmag(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return lib().qo_qmag(_dp(a))


def poly_eval(p, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    out = np.empty(4); lib().qo_poly_eval(ctypes.byref(p.p), _dp(q), _dp(out)); return out


def clamp_infs_to_mean(field):
    """FIELD_3D::clampInfsToMean on a copy; returns (field, mean)."""
    f = np.array(field, dtype=np.float64).reshape(-1).copy()
    m = lib().qo_clamp_infs_to_mean(_dp(f), f.size)
    return f, m


def clamp_nans_to_mean(field):
    """FIELD_3D::clampNansToMean on a copy; returns (field, mean)."""
    f = np.array(field, dtype=np.float64).reshape(-1).copy()
    m = lib().qo_clamp_nans_to_mean(_dp(f), f.size)
    return f, m


def set_borders(grid, field, value):
    """FIELD_3D::setBorders on a copy."""
    f = np.array(field, dtype=np.float64).reshape(-1).copy()
    lib().qo_set_borders(ctypes.byref(grid), _dp(f), float(value))
    return f


def inverted_distance(grid, dist, normalized=False):
    """FIELD_3D::invertedDistance / invertedDistanceNormalized of a signed
    distance field (FIELD_3D.cpp:5241-5352). Returns (field, abs_sum) with
    abs_sum the serial absolute sum of the result (normalized only, else
    None); raises ValueError where the reference would sweep forever."""
    d = np.ascontiguousarray(dist, dtype=np.float64).reshape(-1)
    out = np.empty_like(d)
    a = ctypes.c_double(0.0)
    if normalized:
        st = lib().qo_inverted_distance_normalized(ctypes.byref(grid), _dp(d), _dp(out), ctypes.byref(a))
    else:
        st = lib().qo_inverted_distance(ctypes.byref(grid), _dp(d), _dp(out))
    if st != 0:
        raise ValueError("inverted distance: some cell can never be numbered")
    return out, (a.value if normalized else None)


def weight_score(grid, field, wgrid, weight, alpha, threshold=0.0):
    """OPTIMIZE_3D::weightScore / inverseDistanceScore (OPTIMIZE_3D.cpp:4502-4563)
    and computeMaximumScaledPowerRationalScore (:4484-4498): (score, max_score)."""
    f = np.ascontiguousarray(field, dtype=np.float64).reshape(-1)
    w = np.ascontiguousarray(weight, dtype=np.float64).reshape(-1)
    sc, mx = ctypes.c_double(0.0), ctypes.c_double(0.0)
    lib().qo_weight_score(ctypes.byref(grid), _dp(f), ctypes.byref(wgrid), _dp(w), float(alpha), float(threshold),
                          ctypes.byref(sc), ctypes.byref(mx))
    return sc.value, mx.value


def field_min_max(field):
    """FIELD_3D::fieldMin / fieldMax (FIELD_3D.cpp:2960-2971, 3021-3032): the
    serial scan from cell 0 keeping the first cell that compares smaller /
    larger (pure Python; small fields)."""
    f = [float(v) for v in np.asarray(field, dtype=np.float64).reshape(-1)]
    mn = mx = f[0]
    for v in f:
        if v < mn:
            mn = v
        if v > mx:
            mx = v
    return mn, mx
