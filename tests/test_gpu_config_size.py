"""Oracle checks at the configurations' own sizes, on evenly spaced z-slices:

* the 256^3 rational workload (32 + 32 Gaussian roots, powers U{1,2},
  maxIter 4): the GPU FP64 field and the FP32 + FP64-band field against the
  oracle's FP64 map on 8 z-slices;
* one 256^3 tile of the 1024^3 torus workload (128 + 128 roots, maxIter 3), as
  the bench evaluates it (FP32 + band in the extended range of the
  reference's default build), against the oracle's long double map on 8
  z-slices, and in the FP64 range against the FP64 oracle on 2.

Tolerances: FP64 |dF| <= 1e-9 max(1, |F|) (1e-8 with 256 roots) and the same
non-finite cells; FP32 + band zero sign flips of (float)F > 0, the same
non-finite cells, and the FP64 tolerance inside the band."""
import numpy as np
import pytest

import oracle as O
from quijibo_amd import api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE, MAP_EXTENDED_RANGE

pytestmark = pytest.mark.gpu


def _slices(res, n=8):
    return [int(round(k * (res - 1) / (n - 1))) for k in range(n)]


def _gpu(w, grid, prec, flags):
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags)
    return api.field_eval(grid, r, p, prec, 0.05).cpu().numpy()


def _oracle_slices(w, og, zs, flags):
    t, b = O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers)
    p = O.make_params(w.max_iter, w.escape, w.exp_scaling, w.slice, flags)
    return {z: O.field_map_zrange(og, t, b, p, z, z + 1) for z in zs}


def _compare(name, got, ref, rtol, band=None):
    fin = np.isfinite(ref)
    assert np.array_equal(np.isnan(got), np.isnan(ref)) and np.array_equal(np.isinf(got), np.isinf(ref)), name
    if band is None:
        e = np.abs(got[fin] - ref[fin]) / np.maximum(1.0, np.abs(ref[fin]))
        print("%s: max rel %.2e over %d cells" % (name, e.max(), fin.sum()), flush=True)
        assert e.max() <= rtol
        return
    flips = int(np.count_nonzero((got.astype(np.float32) > 0) != (ref.astype(np.float32) > 0)))
    inb = fin & (np.abs(ref) < band - 0.01)
    e = np.abs(got[inb] - ref[inb]) / np.maximum(1.0, np.abs(ref[inb])) if inb.any() else np.zeros(1)
    d = np.abs(got[fin] - ref[fin])
    print("%s: %d flips, max |dF| %.3g, in-band max rel %.2e over %d cells" % (name, flips, d.max(), e.max(),
                                                                               inb.sum()), flush=True)
    assert flips == 0 and e.max() <= rtol


def test_config2_256_on_eight_slices(gpu):
    w = W.config2(256)
    g = api.make_grid(w.res, w.center, w.lengths)
    og = O.make_grid(w.res, w.center, w.lengths)
    zs = _slices(256)
    ref = _oracle_slices(w, og, zs, O.BLACK_HOLE)
    f64 = _gpu(w, g, FP64, MAP_BLACK_HOLE).reshape(256, -1)
    f32 = _gpu(w, g, FP32_BAND, MAP_BLACK_HOLE).reshape(256, -1)
    for z in zs:
        _compare("c2 z=%d fp64" % z, f64[z], ref[z], 1e-9)
        _compare("c2 z=%d fp32+band" % z, f32[z], ref[z], 1e-9, band=0.05)


def test_config3_tile_on_eight_slices(gpu):
    w = W.config3(1024, 256)
    c, L = W.tile_grid(w, 1, 1, 0, 256)
    res = (256, 256, 256)
    g = api.make_grid(res, c, L)
    og = O.make_grid(res, c, L)
    zs = _slices(256)
    ref = _oracle_slices(w, og, zs, O.BLACK_HOLE | O.EXTENDED)
    f32 = _gpu(w, g, FP32_BAND, MAP_BLACK_HOLE | MAP_EXTENDED_RANGE).reshape(256, -1)
    for z in zs:
        _compare("c3 tile z=%d fp32+band ext" % z, f32[z], ref[z], 1e-8, band=0.05)
    # the FP64 range against the FP64 oracle (NaN and -inf cells included)
    zs2 = zs[2:4]
    ref64 = _oracle_slices(w, og, zs2, O.BLACK_HOLE)
    g64 = _gpu(w, g, FP64, MAP_BLACK_HOLE).reshape(256, -1)
    for z in zs2:
        _compare("c3 tile z=%d fp64" % z, g64[z], ref64[z], 1e-8)
