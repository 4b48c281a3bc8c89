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
from quijibo_amd._lib import FP32_BAND, FP64, MAP_BLACK_HOLE

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
