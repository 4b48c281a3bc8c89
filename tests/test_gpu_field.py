"""GPU parity: the HIP field kernels (through the C ABI) against the oracle.

Tolerances (stated here, used below):
  FP64 kernel: |F_gpu - F_cpu| <= 1e-9 * max(1, |F_cpu|) for finite values
  (1e-8 for the serving workloads with up to 256 roots, whose products
  amplify ulp differences more), and the same non-finite pattern (+inf/-inf/NaN). The kernels use the device math
  library (OCML) and FMA contraction, the oracle glibc without contraction, so
  results differ at the ulp level and the map amplifies that a little.
  FP32 bulk + FP64 band: zero isosurface sign flips ((float)F > 0, the
  marching-cubes predicate) and, inside the band, the FP64 tolerance.
"""
import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_ADAPTIVE, MAP_BLACK_HOLE, MAP_SKIP_ORIGIN

pytestmark = pytest.mark.gpu

RTOL64 = 1e-9


def run_both(w, precision=FP64, flags=MAP_BLACK_HOLE, band=0.05, center=None, lengths=None, rotation=None,
             escape=None, init=None):
    center = center or w.center
    lengths = lengths or w.lengths
    rotation = rotation or (1.0, 0.0, 0.0, 0.0)
    escape = w.escape if escape is None else escape
    g = api.make_grid(w.res, center, lengths, rotation)
    p = api.make_params(w.max_iter, escape, w.exp_scaling, w.slice, flags)
    top = api.POLYNOMIAL_4D(w.top.roots, w.top.powers, w.top.power_scalar)
    bot = api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar)
    r = api.Rational(top, bot)
    out = torch.full((w.voxels,), -7.0, dtype=torch.float64, device="cuda") if init is None else init.clone()
    api.field_eval(g, r, p, precision, band, out=out)
    torch.cuda.synchronize()
    gpu = out.cpu().numpy()
    og = O.make_grid(w.res, center, lengths, rotation)
    oflags = (O.BLACK_HOLE if flags & MAP_BLACK_HOLE else 0) | (O.SKIP_ORIGIN if flags & MAP_SKIP_ORIGIN else 0) | \
             (O.AP_REPASS if flags & MAP_ADAPTIVE else 0)
    cpu = O.field_map(og, O.poly(w.top.roots, w.top.powers, w.top.power_scalar),
                      O.poly(w.bottom.roots, w.bottom.powers, w.bottom.power_scalar),
                      O.make_params(w.max_iter, escape, w.exp_scaling, w.slice, oflags),
                      init=None if init is None else init.cpu().numpy())
    if init is None:
        cpu_written = cpu
    else:
        cpu_written = cpu
    return gpu, cpu_written, g, r


def assert_fp64_close(gpu, cpu, rtol=RTOL64, max_outliers=0):
    fin = np.isfinite(cpu)
    same_class = (np.isfinite(gpu) == fin) & ((np.isnan(gpu) == np.isnan(cpu)))
    bad_class = np.count_nonzero(~same_class)
    err = np.abs(gpu[fin] - cpu[fin]) / np.maximum(1.0, np.abs(cpu[fin]))
    n_out = int(np.count_nonzero(err > rtol)) + bad_class
    print("fp64: n=%d nonfinite=%d class_mismatch=%d max_rel=%.3e outliers=%d" %
          (gpu.size, int((~fin).sum()), bad_class, err.max() if err.size else 0.0, n_out), flush=True)
    assert n_out <= max_outliers
    inf = ~fin & ~np.isnan(cpu)
    assert np.array_equal(gpu[inf], cpu[inf])


def test_fp64_config1_64(gpu):
    w = W.config1(64)
    g, c, _, _ = run_both(w)
    assert_fp64_close(g, c)


def test_fp64_config2_48(gpu):
    w = W.config2(48)
    g, c, _, _ = run_both(w)
    assert_fp64_close(g, c)


def test_fp64_config3_roots_32(gpu):
    w = W.config3(32)
    g, c, _, _ = run_both(w)
    assert_fp64_close(g, c)


def test_fp64_rotated_offcentre_grid(gpu):
    w = W.config2(24)
    w.res = (24, 20, 17)
    q = np.array([0.9, 0.2, -0.3, 0.25])
    q /= np.linalg.norm(q)
    g, c, _, _ = run_both(w, center=(0.1, -0.05, 0.2), lengths=(2.8, 3.1, 2.6), rotation=tuple(q))
    assert_fp64_close(g, c)


def test_fp64_general_exponents_and_scalars(gpu):
    w = W.config2(32)
    rng = np.random.default_rng(5)
    w.top.powers = rng.uniform(0.5, 2.5, size=w.top.n)
    w.bottom.powers = rng.uniform(0.5, 2.5, size=w.bottom.n)
    w.top.power_scalar, w.bottom.power_scalar = 0.9, 1.1
    w.exp_scaling = 0.8
    w.slice = 0.05
    g, c, _, _ = run_both(w)
    assert_fp64_close(g, c)


def signflips(a, b):
    return int(np.count_nonzero((a.astype(np.float32) > 0) != (b.astype(np.float32) > 0)))


@pytest.mark.parametrize("res", [48, 96])
def test_fp32_band_config2(gpu, res):
    w = W.config2(res)
    g, c, _, _ = run_both(w, precision=FP32_BAND, band=0.05)
    flips = signflips(g, c)
    fin = np.isfinite(c) & np.isfinite(g)
    err = np.abs(g[fin] - c[fin])
    # voxels whose FP32 value fell inside the band are FP64-exact; those are at
    # least all voxels with |F| < band - (FP32 error), here taken as 0.01
    inband = np.abs(c) < 0.05 - 0.01
    print("fp32band res=%d flips=%d max|dF|=%.3e band=%d" % (res, flips, err.max(), inband.sum()), flush=True)
    assert flips == 0
    assert np.array_equal(np.isfinite(g), np.isfinite(c))
    e_in = np.abs(g[inband] - c[inband]) / np.maximum(1.0, np.abs(c[inband]))
    assert e_in.size == 0 or e_in.max() <= RTOL64
    assert err.max() < 0.05


def test_fp32_band_config3_roots(gpu):
    w = W.config3(32)
    g, c, _, _ = run_both(w, precision=FP32_BAND, band=0.05)
    assert signflips(g, c) == 0
    assert np.array_equal(np.isnan(g), np.isnan(c))


def test_fp32_band_general_exponents(gpu):
    w = W.config2(40)
    rng = np.random.default_rng(9)
    w.top.powers = rng.uniform(0.5, 2.5, size=w.top.n)
    w.bottom.powers = rng.uniform(0.5, 2.5, size=w.bottom.n)
    g, c, _, _ = run_both(w, precision=FP32_BAND, band=0.05)
    assert signflips(g, c) == 0


def test_ap_mode_overflow_repass(gpu):
    """computeRationalMapAP: escape 20, origin skipped, overflowed cells recomputed with extended range."""
    top = W.RootSet(np.array([[1.9, 0.3, 0.1, 0.0], [-2.1, 0.4, -0.2, 0.0]]), np.array([250.0, 250.0]))
    bot = W.RootSet(np.array([[0.05, -0.1, 0.2, 0.0]]), np.array([3.0]))
    w = W.Workload("ap", (10, 10, 9), top, bot, max_iter=1, escape=20.0, exp_scaling=0.5)
    init = torch.full((w.voxels,), 123.0, dtype=torch.float64, device="cuda")
    g, c, _, _ = run_both(w, flags=MAP_SKIP_ORIGIN | MAP_ADAPTIVE, escape=20.0, init=init)
    assert_fp64_close(g, c, rtol=1e-12)
    assert np.all(np.isfinite(g))


def test_skip_origin_leaves_cell_unwritten(gpu):
    w = W.config1(5)  # odd resolution: the middle cell centre is the origin
    init = torch.full((w.voxels,), 42.0, dtype=torch.float64, device="cuda")
    g, c, _, _ = run_both(w, flags=MAP_SKIP_ORIGIN, init=init, lengths=(3.0, 3.0, 3.0))
    mid = 2 + 5 * 2 + 25 * 2
    assert g[mid] == 42.0 and c[mid] == 42.0
    assert_fp64_close(g, c)


def test_points_eval_matches_oracle(gpu):
    w = W.config5(n_surface=3000, n_exterior=3000)
    pts = np.concatenate([w.extra["surface"], w.extra["exterior"]])
    top = api.POLYNOMIAL_4D(w.top.roots, w.top.powers)
    bot = api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers)
    r = api.Rational(top, bot)
    p = api.make_params(w.max_iter, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE)
    d = torch.from_numpy(pts).cuda()
    f64 = api.points_eval(d, r, p, FP64).cpu().numpy()
    f32 = api.points_eval(d, r, p, FP32_BAND, 0.05).cpu().numpy()
    cpu = O.points_eval(pts, O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers),
                        O.make_params(w.max_iter, 2000.0, 1.0, 0.0, O.BLACK_HOLE))
    assert_fp64_close(f64, cpu)
    assert signflips(f32, cpu) == 0


def test_stats_count_iterations(gpu):
    w = W.config1(16)
    ws = api.Workspace()
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots), api.POLYNOMIAL_4D())
    out = torch.empty(w.voxels, dtype=torch.float64, device="cuda")
    ws.reset_stats()
    api.field_eval(g, r, api.make_params(3, 1e3), FP64, out=out, workspace=ws)
    st = ws.stats()
    _, iters = O.field_map(O.make_grid(w.res, w.center, w.lengths), O.poly(w.top.roots), O.poly(np.zeros((0, 4))),
                           O.make_params(3, 1e3), want_iters=True)
    assert st["voxels"] == w.voxels
    assert st["iterations"] == int(iters.sum())


@pytest.mark.parametrize("k", [0, 1, 4, 8])
def test_serving_workloads_fp64_and_band(gpu, k):
    """Config-4 requests (8-256 roots, maxIter 3-5): many voxels overflow or
    zero out in FP64; the FP32 path must reproduce each FP64 class (-inf from
    |B|^2 overflow, NaN from T overflow) and flip no signs."""
    t, w = W.config4_requests(n=12)[k]
    w.res = (40, 36, 32)
    g64, c, _, _ = run_both(w, precision=FP64)
    assert_fp64_close(g64, c, rtol=1e-8)
    g32, c2, _, _ = run_both(w, precision=FP32_BAND)
    assert signflips(g32, c2) == 0
    assert np.array_equal(np.isnan(g32), np.isnan(c2))
    assert np.array_equal(np.isneginf(g32), np.isneginf(c2)) and np.array_equal(np.isposinf(g32), np.isposinf(c2))
    # band voxels are re-evaluated in FP64 (integer exponents as exact products): FP64 tolerance
    inband = np.isfinite(c2) & (np.abs(c2) < 0.05 - 0.01)
    if inband.any():
        assert np.max(np.abs(g32[inband] - c2[inband]) / np.maximum(1.0, np.abs(c2[inband]))) <= 1e-8


@pytest.mark.parametrize("precision", [FP32_BAND, FP64])
def test_field_eval_replays_in_a_captured_graph(gpu, precision):
    """qj_field_eval enqueues kernels only: captured once in a graph it replays the
    eager result exactly, replay after replay, with other work in between."""
    w = W.config2(40)
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    ws = api.Workspace()
    eager = api.field_eval(g, r, p, precision, 0.05, workspace=ws).clone()  # also sizes the workspace lists
    out = torch.empty_like(eager)
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        api.field_eval(g, r, p, precision, 0.05, out=out, workspace=ws, stream=s)
    torch.cuda.synchronize()
    for k in range(3):
        out.fill_(float(k) - 7.0)
        noise = torch.randn(1 << 16, device="cuda").sum()  # unrelated eager work between replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(torch.nan_to_num(out, 1e300, 1e301, -1e301), torch.nan_to_num(eager, 1e300, 1e301, -1e301))
        assert torch.isnan(out).sum() == torch.isnan(eager).sum()
        del noise


@pytest.mark.parametrize("precision", [FP64, FP32_BAND])
def test_edge_cases(gpu, precision):
    """maxIter 0 and cells already outside the escape radius give log|q0|
    (OPTIMIZE_3D.cpp:1319-1328); zero and negative exponents follow pow's
    guards; degenerate grids and empty point sets."""
    def check(g, c):
        if precision == FP64:
            assert_fp64_close(g, c, rtol=1e-8)
            return
        # FP32 bulk: no sign flips, the same non-finite classes, FP64 values in the band
        assert np.count_nonzero((g.astype(np.float32) > 0) != (c.astype(np.float32) > 0)) == 0
        assert np.array_equal(np.isnan(g), np.isnan(c)) and np.array_equal(np.isinf(g), np.isinf(c))
        inband = np.isfinite(c) & (np.abs(c) < 0.04)
        if inband.any():
            assert np.max(np.abs(g[inband] - c[inband])) <= 1e-8
        # outside the band the FP32 values carry FP32 error, amplified where the map is chaotic

    w = W.config2(12)
    w.max_iter = 0
    g, c, _, _ = run_both(w, precision)
    assert_fp64_close(g, c)  # trivial cells are written by the FP64 pass in both modes
    w = W.config2(12)
    g, c, _, _ = run_both(w, precision, escape=1.2)  # the box corners start outside
    check(g, c)
    outside = np.array([np.sqrt(x * x + y * y + z * z) >= 1.2 for z in (-1.375, 1.375) for y in (-1.375, 1.375)
                        for x in (-1.375, 1.375)])
    assert outside.all()
    w = W.config2(10)
    rng = np.random.default_rng(4)
    w.top.powers = rng.choice([0.0, -1.0, 1.0, 2.0, -0.5], w.top.n)
    w.bottom.powers = rng.choice([0.0, 1.0, -2.0, 1.5], w.bottom.n)
    g, c, _, _ = run_both(w, precision)
    check(g, c)
    for res in ((1, 1, 1), (2, 1, 3), (1, 7, 1)):
        w = W.config1(4)
        w.res = res
        g, c, gr, _ = run_both(w, precision)
        check(g, c)
        v, t = api.marching_cubes(torch.from_numpy(c.copy()).cuda(), gr)
        assert len(v) == 0 and len(t) == 0
    w = W.config1(4)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers))
    empty = torch.empty((0, 3), dtype=torch.float64, device="cuda")
    assert api.points_eval(empty, r, api.make_params(3), precision).numel() == 0


def _special_case_workload():
    """Roots placed on the grid so that the first iteration meets the exact
    forms' special cases: q - r = 0 (a root at a cell centre), v = 0 with
    w < 0 (x, y, z of q - r all zero), and q - r with only its x part zero.
    Voxel q0 = (X, Y, Z, slice) in (w, x, y, z); cell centres of a 4^3 grid on
    [-1.5, 1.5]^3 are +-0.375 and +-1.125."""
    rng = np.random.default_rng(11)
    special_top = np.array([[0.375, -0.375, 1.125, 0.0],    # b = 0 at (0.375, -0.375, 1.125)
                            [0.9, 0.375, 0.375, 0.0],       # v = 0, w = X - 0.9 for Y = Z = 0.375
                            [-0.2, -1.125, 0.3, 0.1]])      # b.x = 0 where Y = -1.125
    special_bot = np.array([[1.4, -0.375, -0.375, 0.0],     # v = 0 (squared) for Y = Z = -0.375
                            [-1.125, 0.375, -1.125, 0.0]])  # b = 0 (squared)
    top = np.vstack([special_top, rng.normal(scale=0.6, size=(5, 4))])
    bot = np.vstack([special_bot, rng.normal(scale=0.6, size=(4, 4))])
    tp = np.array([1.0, 1.0, 2.0, 1.0, 2.0, 1.0, 1.0, 2.0])
    bp = np.array([2.0, 2.0, 1.0, 1.0, 2.0, 1.0])
    return W.Workload("special_cases", (4, 4, 4), W.RootSet(top, tp), W.RootSet(bot, bp), max_iter=3)


def test_exact_form_special_cases_in_the_band(gpu):
    """The FP64 band re-evaluation runs the exact forms without per-term tests
    and repeats a point's product in the guarded form when some |(q - r).x| is
    below 1e-150; roots on cell centres make that happen, and the results keep
    the polar form's special cases (b = 0 gives NaN, v = 0 gives (|w|^e, 0))."""
    w = _special_case_workload()
    g64, c, _, _ = run_both(w, FP64)
    assert_fp64_close(g64, c)
    g32, c32, _, _ = run_both(w, FP32_BAND)
    assert np.array_equal(np.isnan(g32), np.isnan(c)) and np.array_equal(np.isinf(g32), np.isinf(c))
    fin = np.isfinite(c)
    assert np.count_nonzero((g32[fin].astype(np.float32) > 0) != (c[fin].astype(np.float32) > 0)) == 0
    # cells with Y = -0.375 or 0.375 meet (q - r).x = 0 for an exponent-1 root (top roots 0
    # and 1), so the FP32 pass hands them to FP64, whose fast product sees min |b.x| = 0 and
    # runs the guarded form; the FP64 tolerance holds there (data is [z][y][x])
    sp = np.zeros(w.res[::-1], dtype=bool)
    sp[:, 1, :] = True
    sp[:, 2, :] = True
    sp = sp.reshape(-1) & fin
    assert np.max(np.abs(g32[sp] - c[sp]) / np.maximum(1.0, np.abs(c[sp]))) <= 1e-8
    assert np.isnan(c).any()  # the b = 0 cells


@pytest.mark.parametrize("general", [False, True])
def test_fp64_exact_matches_the_oracle_and_the_band(gpu, general):
    """QJ_FP64_EXACT (exponents 1 and 2 as exact products): the FP64 tolerance
    against the oracle's polar form, the same non-finite cells, no sign flips;
    and inside the band the FP32 + FP64 band field is this kernel's value bit
    for bit (the band pass runs the same arithmetic)."""
    from quijibo_amd._lib import FP64_EXACT
    w = W.config2(64)
    if general:
        rng = np.random.default_rng(13)
        w.top.powers = np.where(rng.uniform(size=w.top.n) < 0.3, 1.5, w.top.powers)
    g, c, grid, r = run_both(w, precision=FP64_EXACT)
    assert_fp64_close(g, c)
    assert signflips(g, c) == 0
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    f32 = api.field_eval(grid, r, p, FP32_BAND, 0.05).cpu().numpy()
    inband = np.abs(f32) < 0.04
    assert inband.sum() > 100 and np.array_equal(f32[inband], g[inband])
    # the polar FP64 kernel differs from the exact one in the last bits only
    polar = api.field_eval(grid, r, p, FP64).cpu().numpy()
    fin = np.isfinite(polar)
    assert np.array_equal(np.isfinite(g), fin)
    assert np.max(np.abs(polar[fin] - g[fin]) / np.maximum(1.0, np.abs(polar[fin]))) <= RTOL64
    # points: the same values at the cell centres
    pts = torch.from_numpy(np.stack(np.meshgrid(*[-1.5 + (np.arange(64) + 0.5) * 3.0 / 64] * 3, indexing="ij"),
                                    axis=-1)[:, :, :, ::-1].reshape(-1, 3).copy()).cuda()
    pe = api.points_eval(pts, r, p, FP64_EXACT).cpu().numpy()
    assert np.array_equal(pe, g, equal_nan=True)


def test_sampled_fp64_cross_check_in_the_stats(gpu):
    """qj_workspace_check: the FP32 + band field's error statement, against the
    FP64 kernel at the same evenly spaced cells computed here from its field."""
    w = W.config2(64)
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    ws = api.Workspace()
    f64 = api.field_eval(g, r, p, FP64, workspace=ws).cpu().numpy()
    samples = 5000
    ws.check(samples)
    ws.reset_stats()
    f32 = api.field_eval(g, r, p, FP32_BAND, 0.05, workspace=ws).cpu().numpy()
    st = ws.stats()
    n = w.voxels
    idx = ((2 * np.arange(samples) + 1) * n) // (2 * samples)
    a, b = f32[idx], f64[idx]
    fin = np.isfinite(a) & np.isfinite(b)
    print("check:", {k: st[k] for k in st if k.startswith("check") or k == "checked"}, flush=True)
    assert st["checked"] == samples
    assert st["check_sign_flips"] == signflips(a, b) == 0
    assert st["check_class_mismatch"] == int(np.count_nonzero(np.isfinite(a) != np.isfinite(b)))
    assert abs(st["check_max_abs_dF"] - np.max(np.abs(a[fin] - b[fin]))) <= 1e-12
    assert 0 < st["check_max_abs_dF"] < 0.05
    # off again: no more checked cells
    ws.check(0)
    api.field_eval(g, r, p, FP32_BAND, 0.05, workspace=ws)
    assert ws.stats()["checked"] == samples
