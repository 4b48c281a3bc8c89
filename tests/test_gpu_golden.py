"""The GPU path against the committed vectors (tests/golden/oracle_vectors.npz)."""
import os

import numpy as np
import pytest
import torch

from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE, MAP_LAGRANGE

pytestmark = pytest.mark.gpu
GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.npz"))


def _close(a, b, rtol=1e-9):
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin) and np.array_equal(np.isnan(a), np.isnan(b))
    assert np.array_equal(a[~fin & ~np.isnan(b)], b[~fin & ~np.isnan(b)])
    if fin.any():
        assert np.max(np.abs(a[fin] - b[fin]) / np.maximum(1.0, np.abs(b[fin]))) <= rtol


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_gpu_matches_the_vectors(gpu, name):
    w = W.config1(16) if name == "c1" else W.config2(12)
    res = tuple(int(v) for v in GOLD[name + "_res"])
    g = api.make_grid(res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, MAP_BLACK_HOLE)
    _close(api.field_eval(g, r, p, FP64).cpu().numpy(), GOLD[name + "_field"])
    f32 = api.field_eval(g, r, p, FP32_BAND, 0.05).cpu().numpy()
    ref = GOLD[name + "_field"]
    assert np.count_nonzero((f32.astype(np.float32) > 0) != (ref.astype(np.float32) > 0)) == 0
    # the mesh of the golden field
    v, t = api.marching_cubes(torch.from_numpy(ref.copy()).cuda(), g)
    assert np.array_equal(t, GOLD[name + "_mc_triangles"])
    assert np.array_equal(v, GOLD[name + "_mc_vertices"], equal_nan=True)
    pl = api.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, MAP_BLACK_HOLE | MAP_LAGRANGE)
    _close(api.field_eval(g, r, pl, FP64).cpu().numpy(), GOLD[name + "_lagrange_field"])


def test_gpu_rational_test_known_answer(gpu):
    top = api.POLYNOMIAL_4D(np.array([[1.0, 0, 0, 0], [-1.0, 0, 0, 0]]))
    bot = api.POLYNOMIAL_4D(np.array([[0.0, 0, 0, 0]]), np.array([2.0]))
    r = api.Rational(top, bot)
    pts = torch.tensor([list(GOLD["rational_point"])], dtype=torch.float64, device="cuda")
    F = api.points_eval(pts, r, api.make_params(1, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE), FP64).cpu().numpy()
    assert abs(F[0] - GOLD["rational_log_abs_closed_form"][0]) < 1e-13
