"""The FP32 pass's band list holds a quarter of the cells (at least 2^16, at
most all of them); cells that need FP64 beyond it are marked in the output and
evaluated by the band sweep. With a band wider than every value every cell
needs FP64, so the FP32+band result must be the all-FP64 exact-power result
bit for bit (both re-evaluate with the same FP64 function), with the list
overflowing (3/4 of the cells swept) and not, for grids and points, in the
default and the extended range; and the band statistics must count every
cell once."""
import numpy as np
import pytest
import torch

from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64_EXACT, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE

pytestmark = pytest.mark.gpu


def _setup(w, flags):
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags)
    return r, p


def _bits(t):
    return t.cpu().numpy().view(np.int64)


@pytest.mark.parametrize("res", [40, 64, 96])
@pytest.mark.parametrize("flags", [MAP_BLACK_HOLE, MAP_BLACK_HOLE | MAP_EXTENDED_RANGE])
def test_grid_band_overflow_is_swept(gpu, res, flags):
    w = W.headline(res)
    r, p = _setup(w, flags)
    g = api.make_grid(w.res, w.center, w.lengths)
    ws = api.Workspace()
    n = res ** 3
    ref = api.field_eval(g, r, p, FP64_EXACT, workspace=ws)
    ws.reset_stats()
    got = api.field_eval(g, r, p, FP32_BAND, 1e300, workspace=ws)
    st = ws.stats()
    cap = min(n, max(n // 4, 1 << 16))
    print("res %d: %d cells, list %d, band voxels %d" % (res, n, cap, st["band_voxels"]), flush=True)
    # every cell goes to FP64 except those whose FP32 value is already +-inf
    # (|F| >= the band then holds)
    ninf = int(torch.isinf(ref).sum().item())
    assert n - ninf <= st["band_voxels"] <= n
    assert np.array_equal(_bits(got), _bits(ref))
    # a normal band afterwards: no mark may survive, and the list is not overrun
    ws.reset_stats()
    f = api.field_eval(g, r, p, FP32_BAND, 0.05, workspace=ws)
    st = ws.stats()
    assert 0 < st["band_voxels"] < cap
    assert not np.any(_bits(f) == np.int64(0x7ff40000b0a4d5ee))


@pytest.mark.parametrize("npts", [50000, 300000])
def test_points_band_overflow_is_swept(gpu, npts):
    w = W.config5(n_surface=npts // 2, n_exterior=npts - npts // 2, max_iter=3)
    r, p = _setup(w, MAP_BLACK_HOLE)
    pts = torch.from_numpy(np.concatenate([w.extra["surface"], w.extra["exterior"]])).cuda()
    ws = api.Workspace()
    ref = api.points_eval(pts, r, p, FP64_EXACT, workspace=ws)
    ws.reset_stats()
    got = api.points_eval(pts, r, p, FP32_BAND, 1e300, workspace=ws)
    st = ws.stats()
    ninf = int(torch.isinf(ref).sum().item())
    assert npts - ninf <= st["band_voxels"] <= npts
    assert np.array_equal(_bits(got), _bits(ref))
