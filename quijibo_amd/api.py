"""Python host interface over the C ABI, shaped like the reference classes.

  POLYNOMIAL_4D   roots/powers/power scalar of one polynomial (POLYNOMIAL_4D.h:31-150)
  FIELD_3D        grid geometry + a device-resident float64 tensor, x fastest (FIELD_3D.h:58-80)
  OPTIMIZE_3D     the map members (_maxIterations, _expScaling, _quaternionSlice) and
                  computeLogScaledPowerRationalMap / computeRationalMapAP (OPTIMIZE_3D.h:161-169)
  TRIANGLE_MESH   marching cubes of a FIELD_3D: vertices(), triangles(), writeOBJ (TRIANGLE_MESH.h:42-73)

Device memory and streams come from torch (plumbing only); every computation
runs in libquijibo.so's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import FP32_BAND, FP64, MAP_ADAPTIVE, MAP_BLACK_HOLE, MAP_SKIP_ORIGIN, check

DEFAULT_BAND = 0.05


def _stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def _device_tensor(t, name, numel=None, dtype=torch.float64):
    """Checks a tensor handed to the C ABI as a device pointer: a contiguous CUDA
    tensor of the kernel's element type and, when numel is given, that size.
    The C ABI receives plain pointers, so these checks happen here."""
    if not isinstance(t, torch.Tensor):
        raise ValueError("%s must be a torch tensor" % name)
    if not t.is_cuda:
        raise ValueError("%s must be a CUDA tensor (got device %s)" % (name, t.device))
    if t.dtype != dtype:
        raise ValueError("%s must be %s (got %s)" % (name, dtype, t.dtype))
    if not t.is_contiguous():
        raise ValueError("%s must be contiguous" % name)
    if numel is not None and t.numel() != numel:
        raise ValueError("%s must have %d elements (got %d)" % (name, numel, t.numel()))
    return t


def _cells(grid):
    return int(grid.res[0]) * int(grid.res[1]) * int(grid.res[2])


class POLYNOMIAL_4D:
    """Factored polynomial: prod_i (q - roots[i])^(power_scalar * powers[i])."""

    def __init__(self, roots=None, powers=None, power_scalar=1.0):
        r = np.zeros((0, 4)) if roots is None else np.asarray(roots, dtype=np.float64).reshape(-1, 4)
        self.roots = np.ascontiguousarray(r)
        self.powers = np.ascontiguousarray(np.ones(len(r)) if powers is None
                                           else np.asarray(powers, dtype=np.float64).reshape(-1))
        if self.powers.shape[0] != self.roots.shape[0]:
            raise ValueError("powers must have one entry per root")
        self.power_scalar = float(power_scalar)

    def totalRoots(self):
        return int(self.roots.shape[0])

    def addRoot(self, root, power=1.0):
        self.roots = np.ascontiguousarray(np.vstack([self.roots, np.asarray(root, dtype=np.float64).reshape(1, 4)]))
        self.powers = np.ascontiguousarray(np.append(self.powers, float(power)))

    def c_struct(self):
        return _lib.Poly(self.totalRoots(), self.roots.ctypes.data_as(_lib.c_double_p),
                         self.powers.ctypes.data_as(_lib.c_double_p), self.power_scalar)


class Rational:
    """A top/bottom POLYNOMIAL_4D pair resident in HBM (qj_rational)."""

    def __init__(self, top, bottom=None):
        _lib.load()
        self.top = top
        self.bottom = bottom if bottom is not None else POLYNOMIAL_4D()
        t, b = top.c_struct(), self.bottom.c_struct()
        h = ctypes.c_void_p()
        check(_lib.load().qj_rational_create(ctypes.byref(t), ctypes.byref(b), ctypes.byref(h)))
        self.handle = h

    def update(self, top=None, bottom=None, stream=None):
        if top is not None:
            self.top = top
        if bottom is not None:
            self.bottom = bottom
        t, b = self.top.c_struct(), self.bottom.c_struct()
        check(_lib.load().qj_rational_update(self.handle, ctypes.byref(t), ctypes.byref(b), _stream_ptr(stream)))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().qj_rational_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


class Workspace:
    def __init__(self):
        h = ctypes.c_void_p()
        check(_lib.load().qj_workspace_create(ctypes.byref(h)))
        self.handle = h

    def stats(self, stream=None):
        st = _lib.EvalStats()
        check(_lib.load().qj_workspace_stats(self.handle, _stream_ptr(stream), ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.EvalStats._fields_}

    def reset_stats(self, stream=None):
        check(_lib.load().qj_workspace_reset_stats(self.handle, _stream_ptr(stream)))

    def check(self, samples):
        """Sampled FP64 cross-check of every field evaluated on this workspace
        (qj_workspace_check); the results appear in stats()."""
        check(_lib.load().qj_workspace_check(self.handle, int(samples)))

    def __del__(self):
        try:
            if getattr(self, "handle", None):
                _lib.load().qj_workspace_destroy(self.handle)
                self.handle = None
        except Exception:
            pass


_default_ws = {}


def default_workspace():
    dev = torch.cuda.current_device()
    if dev not in _default_ws:
        _default_ws[dev] = Workspace()
    return _default_ws[dev]


def make_grid(res, center=(0.0, 0.0, 0.0), lengths=(3.0, 3.0, 3.0), rotation=(1.0, 0.0, 0.0, 0.0)):
    if isinstance(res, int):
        res = (res, res, res)
    g = _lib.Grid()
    g.res[:] = [int(v) for v in res]
    g.center[:] = [float(v) for v in center]
    g.lengths[:] = [float(v) for v in lengths]
    g.rotation[:] = [float(v) for v in rotation]
    return g


def make_params(max_iter=3, escape=2000.0, exp_scaling=1.0, slice_=0.0, flags=MAP_BLACK_HOLE):
    return _lib.MapParams(int(max_iter), float(escape), float(exp_scaling), float(slice_), int(flags))


class FIELD_3D:
    """Cell-centred grid with a float64 device tensor, data[z][y][x]."""

    def __init__(self, xRes, yRes, zRes, center=(0.0, 0.0, 0.0), lengths=(1.0, 1.0, 1.0), data=None,
                 rotation=(1.0, 0.0, 0.0, 0.0), device=None):
        self.res = (int(xRes), int(yRes), int(zRes))
        self.center = tuple(float(v) for v in center)
        self.lengths = tuple(float(v) for v in lengths)
        self.rotation = tuple(float(v) for v in rotation)
        dev = device if device is not None else torch.device("cuda", torch.cuda.current_device())
        if data is None:
            self.data = torch.zeros((self.res[2], self.res[1], self.res[0]), dtype=torch.float64, device=dev)
        else:
            if data.dtype != torch.float64 or not data.is_cuda or not data.is_contiguous():
                raise ValueError("FIELD_3D data must be a contiguous float64 CUDA tensor")
            self.data = data.view(self.res[2], self.res[1], self.res[0])

    def xRes(self):
        return self.res[0]

    def yRes(self):
        return self.res[1]

    def zRes(self):
        return self.res[2]

    def totalCells(self):
        return self.res[0] * self.res[1] * self.res[2]

    def dx(self):
        return self.lengths[0] / self.res[0]

    def grid(self):
        return make_grid(self.res, self.center, self.lengths, self.rotation)

    def numpy(self):
        return self.data.detach().cpu().numpy().reshape(-1)

    def clampInfsToMean(self, workspace=None, stream=None, return_mean=True):
        """FIELD_3D::clampInfsToMean (FIELD_3D.cpp:5182-5207) on the device; returns the mean used
        (which waits for the stream), or None without waiting when return_mean is False."""
        return _clamp(self, "qj_field_clamp_infs_to_mean", workspace, stream, return_mean)

    def clampNansToMean(self, workspace=None, stream=None, return_mean=True):
        """FIELD_3D::clampNansToMean (FIELD_3D.cpp:5212-5234) on the device; returns the mean used
        (which waits for the stream), or None without waiting when return_mean is False."""
        return _clamp(self, "qj_field_clamp_nans_to_mean", workspace, stream, return_mean)

    def setBorders(self, value, stream=None):
        """FIELD_3D::setBorders (FIELD_3D.cpp:5357-5382): the six faces take value."""
        g = self.grid()
        check(_lib.load().qj_field_set_borders(ctypes.byref(g), float(value), _ptr(self.data), _stream_ptr(stream)))

    def fieldMin(self, workspace=None, stream=None):
        """FIELD_3D::fieldMin (FIELD_3D.cpp:3021-3032) on the device."""
        return field_min_max(self.data.view(-1), workspace, stream)[0]

    def fieldMax(self, workspace=None, stream=None):
        """FIELD_3D::fieldMax (FIELD_3D.cpp:2960-2971) on the device."""
        return field_min_max(self.data.view(-1), workspace, stream)[1]

    def invertedDistance(self, workspace=None, stream=None):
        """FIELD_3D::invertedDistance (FIELD_3D.cpp:5241-5297) of this signed
        distance field, as a new FIELD_3D on the same grid."""
        out, _ = inverted_distance(self.data.view(-1), self.grid(), False, workspace, stream)
        return FIELD_3D(*self.res, center=self.center, lengths=self.lengths, data=out, rotation=self.rotation)

    def invertedDistanceNormalized(self, workspace=None, stream=None):
        """FIELD_3D::invertedDistanceNormalized (FIELD_3D.cpp:5316-5352):
        (field, absolute sum of the result)."""
        out, a = inverted_distance(self.data.view(-1), self.grid(), True, workspace, stream)
        return FIELD_3D(*self.res, center=self.center, lengths=self.lengths, data=out, rotation=self.rotation), a


def _clamp(field, fn, workspace, stream, return_mean=True):
    _device_tensor(field.data, "field", field.totalCells())
    ws = workspace or default_workspace()
    mean = ctypes.c_double()
    check(getattr(_lib.load(), fn)(ws.handle, _ptr(field.data), field.totalCells(),
                                   ctypes.byref(mean) if return_mean else None, _stream_ptr(stream)))
    return mean.value if return_mean else None


def field_min_max(field, workspace=None, stream=None):
    """qj_field_min_max: (fieldMin, fieldMax) of a float64 CUDA tensor with the
    reference's serial-scan semantics (FIELD_3D.cpp:2960-3032)."""
    n = field.numel()
    _device_tensor(field, "field", n)
    if n < 1:
        raise ValueError("field_min_max: empty field")
    ws = workspace or default_workspace()
    mn, mx = ctypes.c_double(), ctypes.c_double()
    check(_lib.load().qj_field_min_max(ws.handle, _ptr(field), n, ctypes.byref(mn), ctypes.byref(mx),
                                       _stream_ptr(stream)))
    return mn.value, mx.value


def inverted_distance(dist, grid, normalized=False, workspace=None, stream=None, out=None):
    """qj_field_inverted_distance: FIELD_3D::invertedDistance(Normalized) of a
    signed distance field (float64 CUDA tensor on `grid`). Returns (out, abs
    sum of out)."""
    n = _cells(grid)
    _device_tensor(dist, "dist", n)
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=dist.device)
    _device_tensor(out, "out", n)
    ws = workspace or default_workspace()
    a = ctypes.c_double()
    check(_lib.load().qj_field_inverted_distance(ws.handle, ctypes.byref(grid), _ptr(dist), 1 if normalized else 0,
                                                 _ptr(out), ctypes.byref(a), _stream_ptr(stream)))
    return out, a.value


def weight_scores(field, grid, weight, wgrid, alpha, thresholds=(0.0,), workspace=None, stream=None):
    """qj_weight_scores: OPTIMIZE_3D::weightScore / inverseDistanceScore for each
    threshold and computeMaximumScaledPowerRationalScore (OPTIMIZE_3D.cpp:4484-4563).
    Returns (scores as a numpy array, max_score)."""
    _device_tensor(field, "field", _cells(grid))
    _device_tensor(weight, "weight", _cells(wgrid))
    t = np.ascontiguousarray(thresholds, dtype=np.float64).reshape(-1)
    sc = np.zeros(t.size, dtype=np.float64)
    mx = ctypes.c_double()
    ws = workspace or default_workspace()
    check(_lib.load().qj_weight_scores(ws.handle, ctypes.byref(grid), _ptr(field), ctypes.byref(wgrid), _ptr(weight),
                                       float(alpha), t.ctypes.data_as(_lib.c_double_p), int(t.size),
                                       sc.ctypes.data_as(_lib.c_double_p), ctypes.byref(mx), _stream_ptr(stream)))
    return sc, mx.value


def field_eval(grid, rational, params, precision=FP64, band=DEFAULT_BAND, out=None, workspace=None, stream=None):
    """qj_field_eval into a float64 CUDA tensor (allocated when out is None)."""
    n = _cells(grid)
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device="cuda")
    _device_tensor(out, "out", n)
    ws = workspace or default_workspace()
    check(_lib.load().qj_field_eval(ws.handle, ctypes.byref(grid), rational.handle, ctypes.byref(params),
                                    int(precision), float(band), _ptr(out), _stream_ptr(stream)))
    return out


def points_eval(points, rational, params, precision=FP64, band=DEFAULT_BAND, out=None, workspace=None, stream=None):
    if not isinstance(points, torch.Tensor) or points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must be an (n, 3) float64 CUDA tensor")
    pts = _device_tensor(points.contiguous(), "points")
    n = pts.shape[0]
    if out is None:
        out = torch.empty(n, dtype=torch.float64, device=pts.device)
    _device_tensor(out, "out", n)
    ws = workspace or default_workspace()
    check(_lib.load().qj_points_eval(ws.handle, _ptr(pts), n, rational.handle, ctypes.byref(params),
                                     int(precision), float(band), _ptr(out), _stream_ptr(stream)))
    return out


def marching_cubes(field, grid, iso=0.0, workspace=None, stream=None, device_output=False):
    """qj_mc_run + copy: returns (vertices (nv,3) float64, triangles (nt,3) int32)."""
    _device_tensor(field, "field", _cells(grid))
    ws = workspace or default_workspace()
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = _lib.load()
    check(L.qj_mc_run(ws.handle, _ptr(field), ctypes.byref(grid), float(iso), ctypes.byref(nv), ctypes.byref(nt),
                      _stream_ptr(stream)))
    if device_output:
        v = torch.empty((nv.value, 3), dtype=torch.float64, device=field.device)
        t = torch.empty((nt.value, 3), dtype=torch.int32, device=field.device)
        check(L.qj_mc_copy(ws.handle, _ptr(v) if nv.value else None, _ptr(t) if nt.value else None,
                           _stream_ptr(stream)))
        return v, t
    v = np.empty((nv.value, 3), dtype=np.float64)
    t = np.empty((nt.value, 3), dtype=np.int32)
    check(L.qj_mc_copy(ws.handle, v.ctypes.data_as(ctypes.c_void_p) if nv.value else None,
                       t.ctypes.data_as(ctypes.c_void_p) if nt.value else None, _stream_ptr(stream)))
    return v, t


def marching_cubes_tiles(fields, grids, iso=0.0, workspace=None, stream=None, out=None):
    """qj_mc_run_tiles: marching cubes of several fields with one read-back of
    the counts. Returns (vertices, triangles, counts) with the meshes one after
    another (triangle indices local to their field) as device tensors and the
    per-field (vertices, triangles) counts. out = (vertices (cap_v, 3) float64,
    triangles (cap_t, 3) int32) CUDA tensors: the meshes are written straight
    into them (qj_mc_run_tiles_into, no copy) and views of their first rows are
    returned; QJError (status QJ_ERR_CAPACITY) when they are too small."""
    n = len(fields)
    if len(grids) != n:
        raise ValueError("one grid per field")
    for k, (f, g) in enumerate(zip(fields, grids)):
        _device_tensor(f, "fields[%d]" % k, _cells(g))
    ws = workspace or default_workspace()
    ptrs = (ctypes.c_void_p * max(n, 1))(*[f.data_ptr() for f in fields])
    garr = (_lib.Grid * max(n, 1))(*grids)
    nv = (ctypes.c_int64 * max(n, 1))()
    nt = (ctypes.c_int64 * max(n, 1))()
    L = _lib.load()
    if out is not None:
        ov, ot = out
        for name, t, dt in (("out vertices", ov, torch.float64), ("out triangles", ot, torch.int32)):
            if (not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != dt or t.dim() != 2 or t.shape[1] != 3
                    or not t.is_contiguous()):
                raise ValueError("%s must be a contiguous (n, 3) %s CUDA tensor" % (name, dt))
        check(L.qj_mc_run_tiles_into(ws.handle, n, ctypes.cast(ptrs, ctypes.c_void_p), garr, float(iso), _ptr(ov),
                                     int(ov.shape[0]), _ptr(ot), int(ot.shape[0]), nv, nt, _stream_ptr(stream)))
        counts = [(nv[k], nt[k]) for k in range(n)]
        tv, tt = sum(c[0] for c in counts), sum(c[1] for c in counts)
        return ov[:tv], ot[:tt], counts
    check(L.qj_mc_run_tiles(ws.handle, n, ctypes.cast(ptrs, ctypes.c_void_p), garr, float(iso), nv, nt,
                            _stream_ptr(stream)))
    counts = [(nv[k], nt[k]) for k in range(n)]
    tv, tt = sum(c[0] for c in counts), sum(c[1] for c in counts)
    dev = fields[0].device if n else torch.device("cuda")
    v = torch.empty((tv, 3), dtype=torch.float64, device=dev)
    t = torch.empty((tt, 3), dtype=torch.int32, device=dev)
    check(L.qj_mc_copy(ws.handle, _ptr(v) if tv else None, _ptr(t) if tt else None, _stream_ptr(stream)))
    return v, t, counts


class OPTIMIZE_3D:
    """The map-evaluation members of OPTIMIZE_3D (OPTIMIZE_3D.h:67-81, 140-169)."""

    def __init__(self, top, bottom=None, maxIterations=3, expScaling=1.0, quaternionSlice=0.0):
        self._top = top
        self._bottom = bottom if bottom is not None else POLYNOMIAL_4D()
        self._maxIterations = int(maxIterations)
        self._expScaling = float(expScaling)
        self._quaternionSlice = float(quaternionSlice)
        self._rational = None

    def _sync_rational(self, top, bottom):
        if (self._rational is None or self._rational.top is not top or self._rational.bottom is not bottom
                or self._rational.top.totalRoots() != top.totalRoots()
                or self._rational.bottom.totalRoots() != bottom.totalRoots()):
            self._rational = Rational(top, bottom)
        else:
            self._rational.update(top, bottom)
        return self._rational

    def computeLogScaledPowerRationalMap(self, top=None, bottom=None, fractal=None, verbose=False,
                                         precision=FP64, band=DEFAULT_BAND):
        """OPTIMIZE_3D.cpp:1272-1370: escape 2000, black-hole stop, field = log|q|."""
        top = top or self._top
        bottom = bottom or self._bottom
        r = self._sync_rational(top, bottom)
        p = make_params(self._maxIterations, 2000.0, self._expScaling, self._quaternionSlice, MAP_BLACK_HOLE)
        field_eval(fractal.grid(), r, p, precision, band, out=fractal.data.view(-1))
        return fractal

    def computeRationalMapAP(self, top=None, bottom=None, fractal=None, verbose=False):
        """OPTIMIZE_3D.cpp:1374-1568: escape 20, origin cells skipped, overflow re-pass."""
        top = top or self._top
        bottom = bottom or self._bottom
        r = self._sync_rational(top, bottom)
        p = make_params(self._maxIterations, 20.0, self._expScaling, self._quaternionSlice,
                        MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
        field_eval(fractal.grid(), r, p, FP64, 0.0, out=fractal.data.view(-1))
        return fractal


class TRIANGLE_MESH:
    """TRIANGLE_MESH(const FIELD_3D&): staged marching cubes on the device."""

    def __init__(self, field, iso=0.0):
        self._vertices, self._triangles = marching_cubes(field.data.view(-1), field.grid(), iso)

    def vertices(self):
        return self._vertices

    def triangles(self):
        return self._triangles

    def writeOBJ(self, filename):
        """TRIANGLE_MESH::writeOBJ (TRIANGLE_MESH.cpp:2762-2800): "v %.16f", 1-based "f i j k"."""
        with open(filename, "w") as f:
            for v in self._vertices:
                f.write("v %.16f %.16f %.16f\n" % (v[0], v[1], v[2]))
            for t in self._triangles:
                f.write("f %i %i %i\n" % (t[0] + 1, t[1] + 1, t[2] + 1))
        return True


def ap_score_gradients(grid, rational, params, wgrid, weight, tanh_alpha=10.0, tanh_threshold=0.0, field=None,
                       workspace=None, stream=None):
    """One-iteration AP score and exponent gradients (qj_ap_score_gradients).

    weight: float64 CUDA tensor on wgrid (x fastest). field: float64 CUDA tensor
    that receives the AP field (cells at the origin are left as they are).
    Returns a dict with score, grad_exp, grad_top, grad_bottom, grad_top_scalar,
    grad_bottom_scalar, adaptive_cells and the field tensor."""
    n = _cells(grid)
    if field is None:
        field = torch.zeros(n, dtype=torch.float64, device="cuda")
    _device_tensor(field, "field", n)
    w = _device_tensor(weight.contiguous() if isinstance(weight, torch.Tensor) else weight, "weight", _cells(wgrid))
    nt, nb = rational.top.totalRoots(), rational.bottom.totalRoots()
    gt = np.zeros(max(nt, 1))
    gb = np.zeros(max(nb, 1))
    out = _lib.APGrads()
    ws = workspace or default_workspace()
    check(_lib.load().qj_ap_score_gradients(
        ws.handle, ctypes.byref(grid), rational.handle, ctypes.byref(params), ctypes.byref(wgrid), _ptr(w),
        float(tanh_alpha), float(tanh_threshold), _ptr(field), gt.ctypes.data_as(ctypes.c_void_p),
        gb.ctypes.data_as(ctypes.c_void_p), ctypes.byref(out), _stream_ptr(stream)))
    return dict(score=out.score, grad_exp=out.grad_exp, grad_top_scalar=out.grad_top_scalar,
                grad_bottom_scalar=out.grad_bottom_scalar, adaptive_cells=out.adaptive_cells,
                grad_top=gt[:nt].copy(), grad_bottom=gb[:nb].copy(), field=field)


def points_adjoint(points, rational, params, dJdF, precision=FP64, workspace=None, stream=None):
    """qj_points_adjoint: dJ/d(roots) ((nt+nb) x 4, top first) and dJ/d(powers) (nt+nb)."""
    if not isinstance(points, torch.Tensor) or points.dim() != 2 or points.shape[1] != 3:
        raise ValueError("points must be an (n, 3) float64 CUDA tensor")
    pts = _device_tensor(points.contiguous(), "points")
    if not isinstance(dJdF, torch.Tensor):
        raise ValueError("dJdF must be a torch tensor")
    g = _device_tensor(dJdF.contiguous().to(torch.float64), "dJdF", pts.shape[0])
    nr = rational.top.totalRoots() + rational.bottom.totalRoots()
    gr = torch.empty((nr, 4), dtype=torch.float64, device=pts.device)
    gp = torch.empty(nr, dtype=torch.float64, device=pts.device)
    ws = workspace or default_workspace()
    check(_lib.load().qj_points_adjoint(ws.handle, _ptr(pts), pts.shape[0], rational.handle, ctypes.byref(params),
                                        int(precision), _ptr(g), _ptr(gr), _ptr(gp), _stream_ptr(stream)))
    return gr, gp


def nonlinear_mesh(grid, rational, params, iso=0.0, precision=FP32_BAND, band=DEFAULT_BAND, chunk_slices=64,
                   workspace=None, stream=None, allow_inconsistent=False, device_output=False):
    """qj_nlmc_run + copy: the non-linear low-memory mesher. Returns (vertices, triangles), as host
    arrays, or as device tensors with device_output=True."""
    ws = workspace or default_workspace()
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = _lib.load()
    st = L.qj_nlmc_run(ws.handle, ctypes.byref(grid), rational.handle, ctypes.byref(params), float(iso),
                       int(precision), float(band), int(chunk_slices), ctypes.byref(nv), ctypes.byref(nt),
                       _stream_ptr(stream))
    if st != 0 and not (allow_inconsistent and st == 5):
        check(st)
    if device_output:
        v = torch.empty((nv.value, 3), dtype=torch.float64, device="cuda")
        t = torch.empty((nt.value, 3), dtype=torch.int32, device="cuda")
        check(L.qj_mc_copy(ws.handle, _ptr(v) if nv.value else None, _ptr(t) if nt.value else None,
                           _stream_ptr(stream)))
        return v, t
    v = np.empty((nv.value, 3), dtype=np.float64)
    t = np.empty((nt.value, 3), dtype=np.int32)
    check(L.qj_mc_copy(ws.handle, v.ctypes.data_as(ctypes.c_void_p) if nv.value else None,
                       t.ctypes.data_as(ctypes.c_void_p) if nt.value else None, _stream_ptr(stream)))
    return v, t


def nonlinear_mesh_field(field, grid, rational, params, iso=0.0, workspace=None, stream=None, allow_inconsistent=False):
    """qj_nlmc_field: the non-linear mesher on a given field (cases from its
    signs, vertices by bisection on the map). Returns (vertices, triangles)."""
    _device_tensor(field, "field", _cells(grid))
    ws = workspace or default_workspace()
    nv = ctypes.c_int64()
    nt = ctypes.c_int64()
    L = _lib.load()
    st = L.qj_nlmc_field(ws.handle, _ptr(field), ctypes.byref(grid), rational.handle, ctypes.byref(params),
                         float(iso), ctypes.byref(nv), ctypes.byref(nt), _stream_ptr(stream))
    if st != 0 and not (allow_inconsistent and st == 5):
        check(st)
    v = np.empty((nv.value, 3), dtype=np.float64)
    t = np.empty((nt.value, 3), dtype=np.int32)
    check(L.qj_mc_copy(ws.handle, v.ctypes.data_as(ctypes.c_void_p) if nv.value else None,
                       t.ctypes.data_as(ctypes.c_void_p) if nt.value else None, _stream_ptr(stream)))
    return v, t


def nonlinear_mesh_cache(workspace=None, stream=None):
    """qj_nlmc_cache: (edges n x 4 int32 [x, y, z, axis], cells m x 4 int32 [x, y, z, flag]) of the last
    non-linear mesh on the workspace."""
    ws = workspace or default_workspace()
    L = _lib.load()
    ne, nc = ctypes.c_int64(), ctypes.c_int64()
    check(L.qj_nlmc_cache(ws.handle, ctypes.byref(ne), None, ctypes.byref(nc), None, _stream_ptr(stream)))
    e = np.zeros((ne.value, 4), dtype=np.int32)
    c = np.zeros((nc.value, 4), dtype=np.int32)
    check(L.qj_nlmc_cache(ws.handle, None, e.ctypes.data_as(ctypes.c_void_p) if ne.value else None, None,
                          c.ctypes.data_as(ctypes.c_void_p) if nc.value else None, _stream_ptr(stream)))
    return e, c


def nonlinear_mesh_resume(grid, rational, params, edges, cells, iso=0.0, workspace=None, stream=None):
    """qj_nlmc_resume: the mesh rebuilt from an edge cache; returns (vertices, triangles)."""
    ws = workspace or default_workspace()
    L = _lib.load()
    e = np.ascontiguousarray(edges, dtype=np.int32).reshape(-1, 4)
    c = np.ascontiguousarray(cells, dtype=np.int32).reshape(-1, 4)
    nv, nt = ctypes.c_int64(), ctypes.c_int64()
    check(L.qj_nlmc_resume(ws.handle, ctypes.byref(grid), rational.handle, ctypes.byref(params), float(iso),
                           e.shape[0], e.ctypes.data_as(ctypes.c_void_p) if e.shape[0] else None, c.shape[0],
                           c.ctypes.data_as(ctypes.c_void_p) if c.shape[0] else None, ctypes.byref(nv),
                           ctypes.byref(nt), _stream_ptr(stream)))
    v = np.empty((nv.value, 3), dtype=np.float64)
    t = np.empty((nt.value, 3), dtype=np.int32)
    check(L.qj_mc_copy(ws.handle, v.ctypes.data_as(ctypes.c_void_p) if nv.value else None,
                       t.ctypes.data_as(ctypes.c_void_p) if nt.value else None, _stream_ptr(stream)))
    return v, t
