"""GPU marching cubes (through the C ABI) against the oracle mesher on the same
field: vertex count, triangle count, vertex coordinates and triangle indices
must be bit-identical."""
import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP64, MAP_BLACK_HOLE, QJError

pytestmark = pytest.mark.gpu


def mc_both(field_np, res, center=(0, 0, 0), lengths=(3, 3, 3), rotation=(1, 0, 0, 0), iso=0.0):
    g = api.make_grid(res, center, lengths, rotation)
    d = torch.from_numpy(np.ascontiguousarray(field_np)).cuda()
    v, t = api.marching_cubes(d, g, iso)
    og = O.make_grid(res, center, lengths, rotation)
    f = field_np if iso == 0.0 else field_np - iso
    ov, ot = O.marching_cubes(og, f)
    return v, t, ov, ot


def assert_same_mesh(v, t, ov, ot):
    print("mc: gpu %d verts %d tris, cpu %d verts %d tris" % (len(v), len(t), len(ov), len(ot)), flush=True)
    assert v.shape == ov.shape and t.shape == ot.shape
    assert np.array_equal(t, ot)
    same = (v == ov) | (np.isnan(v) & np.isnan(ov))
    assert same.all(), np.argwhere(~same)[:5]


def test_sphere(gpu):
    n = 40
    ax = -1.5 + (np.arange(n) + 0.5) * 3.0 / n
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    f = 0.9 - np.sqrt(X ** 2 + (Y - 0.1) ** 2 + Z ** 2)
    assert_same_mesh(*mc_both(f.reshape(-1), (n, n, n)))


@pytest.mark.parametrize("flag", [1, 97, 158, 255, 0, 105, 60])
def test_single_cells(gpu, flag):
    f = np.empty(8)
    for z in range(2):
        for y in range(2):
            for x in range(2):
                bit = 4 * x + 2 * y + z
                f[x + 2 * y + 4 * z] = (1.0 if (flag >> bit) & 1 else -1.0) * (1 + 0.1 * bit)
    assert_same_mesh(*mc_both(f, (2, 2, 2), lengths=(2, 2, 2)))


def test_random_field_all_cases(gpu):
    rng = np.random.default_rng(2)
    res = (23, 17, 13)
    f = rng.normal(size=res[0] * res[1] * res[2])
    f[rng.integers(0, f.size, 40)] = np.nan
    f[rng.integers(0, f.size, 40)] = -np.inf
    f[rng.integers(0, f.size, 40)] = 1e-50
    v, t, ov, ot = mc_both(f, res, center=(0.2, 0.1, -0.3), lengths=(2.0, 1.5, 1.1),
                           rotation=(0.8, 0.36, 0.48, 0.0))
    assert_same_mesh(v, t, ov, ot)


def test_exact_iso_values_and_degenerate_triangles(gpu):
    """Values exactly at the isovalue put vertices on grid points, so triangles
    with coincident corners occur and the 1e-7 degenerate filter drops them
    (TRIANGLE_MESH.cpp:1088-1102). Once with a rotated grid (the kernel skips
    the distance test for triangles with fewer than two vertices near an edge
    end) and once with a spacing below 1e-6 (every triangle takes the test)."""
    rng = np.random.default_rng(5)
    res = (19, 16, 14)
    f = rng.integers(-1, 2, size=res[0] * res[1] * res[2]).astype(np.float64)
    v, t, ov, ot = mc_both(f, res, center=(0.1, -0.2, 0.3), lengths=(2.0, 1.7, 1.5), rotation=(0.8, 0.36, 0.48, 0.0))
    assert_same_mesh(v, t, ov, ot)
    assert len(np.unique(v, axis=0)) < len(v)  # coincident vertices: the filter had cases to decide
    v, t, ov, ot = mc_both(f, res, lengths=(1e-5, 1e-5, 1e-5))
    assert_same_mesh(v, t, ov, ot)


def test_measure_list_overflow_is_swept(gpu):
    """A spacing below 1e-6 sends every active cell to the FP64 degenerate
    test. On 64^3 random values nearly every cell is active, far more than the
    measure list's 251 sub-lists hold (a sixteenth of the points, at least 251
    x 256 entries), so most cells are marked in their mask byte and measured by
    the sweep instead: the mesh must still be the oracle's."""
    rng = np.random.default_rng(17)
    res = (64, 64, 64)
    f = rng.normal(size=res[0] * res[1] * res[2])
    f[rng.integers(0, f.size, 200)] = 0.0  # vertices on grid points: coincident corners to drop
    v, t, ov, ot = mc_both(f, res, lengths=(1e-5, 1e-5, 1e-5))
    assert_same_mesh(v, t, ov, ot)
    assert len(t) > 100000


def test_iso_shift(gpu):
    rng = np.random.default_rng(3)
    res = (12, 12, 12)
    f = rng.normal(size=12 ** 3)
    assert_same_mesh(*mc_both(f, res, iso=0.3))


@pytest.mark.parametrize("cfg", ["c1", "c3"])
def test_julia_fields(gpu, cfg):
    w = W.config1(64) if cfg == "c1" else W.config3(48)
    g = api.make_grid(w.res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    out = api.field_eval(g, r, api.make_params(w.max_iter, w.escape, 1.0, 0.0, MAP_BLACK_HOLE), FP64)
    f = out.cpu().numpy()
    v, t, ov, ot = mc_both(f, w.res)
    assert len(t) > 1000
    assert_same_mesh(v, t, ov, ot)


def test_c3_full_tile_mesh_is_bit_exact(gpu):
    """One 256^3 tile of the 1024^3 torus workload (128+128 roots, maxIter 3), FP32+band
    field: the GPU mesh equals the oracle's marching cubes on the same field."""
    from quijibo_amd._lib import FP32_BAND
    w = W.config3(1024, 256)
    c, tl = W.tile_grid(w, 1, 1, 0, 256)
    g = api.make_grid((256, 256, 256), c, tl)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    out = api.field_eval(g, r, api.make_params(w.max_iter, w.escape, 1.0, 0.0, MAP_BLACK_HOLE), FP32_BAND, 0.05)
    f = out.cpu().numpy()
    v, t = api.marching_cubes(out, g)
    ov, ot = O.marching_cubes(O.make_grid((256, 256, 256), c, tl), f)
    print("c3 tile: %d vertices %d triangles" % (len(v), len(t)), flush=True)
    assert len(t) > 10000
    assert_same_mesh(v, t, ov, ot)  # vertices on edges with an infinite end are NaN on both sides


def test_tiles_in_one_call_match_single_calls_and_the_oracle(gpu):
    """qj_mc_run_tiles: several fields of different sizes meshed with one
    read-back; each field's part equals qj_mc_run on it alone and the oracle's
    mesh. A fresh workspace whose first guess of the buffer sizes is too small
    (random signs: about 1.5 vertices per point) meshes again at the right size."""
    rng = np.random.default_rng(4)
    specs = [((33, 29, 31), (0.1, 0.0, -0.2), (2.0, 1.8, 1.9)), ((40, 40, 40), (0, 0, 0), (3, 3, 3)),
             ((17, 23, 9), (0.3, -0.1, 0.0), (1.0, 1.5, 0.7)), ((2, 2, 2), (0, 0, 0), (1, 1, 1))]
    fields, grids, hosts = [], [], []
    for k, (res, c, L) in enumerate(specs):
        if k == 1:
            ax = -1.5 + (np.arange(40) + 0.5) * 3.0 / 40
            Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
            f = (0.9 - np.sqrt(X ** 2 + (Y - 0.1) ** 2 + Z ** 2)).reshape(-1)
        else:
            f = rng.normal(size=int(np.prod(res)))
        hosts.append(f)
        fields.append(torch.from_numpy(np.ascontiguousarray(f)).cuda())
        grids.append(api.make_grid(res, c, L))
    ws = api.Workspace()
    v, t, counts = api.marching_cubes_tiles(fields, grids, workspace=ws)
    v, t = v.cpu().numpy(), t.cpu().numpy()
    assert counts[0][0] > np.prod(specs[0][0]) // 4  # more vertices than the first guess
    ov0 = ot0 = 0
    for k, (res, c, L) in enumerate(specs):
        nv, nt = counts[k]
        sv, st = api.marching_cubes(fields[k], grids[k], workspace=api.Workspace())
        assert np.array_equal(v[ov0:ov0 + nv], sv) and np.array_equal(t[ot0:ot0 + nt], st), k
        ov, ot = O.marching_cubes(O.make_grid(res, c, L), hosts[k])
        assert_same_mesh(sv, st, ov, ot)
        ov0 += nv
        ot0 += nt
    assert ov0 == len(v) and ot0 == len(t)
    # the same workspace again: buffers now large enough, one pass
    v2, t2, counts2 = api.marching_cubes_tiles(fields, grids, workspace=ws)
    assert counts2 == counts and torch.equal(v2.cpu(), torch.from_numpy(v)) and np.array_equal(t2.cpu().numpy(), t)
    # straight into caller buffers (qj_mc_run_tiles_into): larger than needed, the same mesh in their
    # first rows and the tails untouched; too small, a capacity error that reports the sizes needed
    bv = torch.full((len(v) + 7, 3), -5.0, dtype=torch.float64, device="cuda")
    bt = torch.full((len(t) + 5, 3), -9, dtype=torch.int32, device="cuda")
    v3, t3, counts3 = api.marching_cubes_tiles(fields, grids, workspace=api.Workspace(), out=(bv, bt))
    assert counts3 == counts and v3.data_ptr() == bv.data_ptr()
    assert np.array_equal(v3.cpu().numpy(), v) and np.array_equal(t3.cpu().numpy(), t)
    assert bool((bv[len(v):] == -5.0).all()) and bool((bt[len(t):] == -9).all())
    small = (torch.empty((len(v) - 1, 3), dtype=torch.float64, device="cuda"),
             torch.empty((len(t), 3), dtype=torch.int32, device="cuda"))
    with pytest.raises(QJError) as e:
        api.marching_cubes_tiles(fields, grids, workspace=ws, out=small)
    assert e.value.status == 4 and str(len(v)) in str(e.value)
