"""qj_field_eval_noalloc (the allocation-free field entry point) through the C
ABI against qj_field_eval on the same inputs: bit-identical fields on two
calls in a row, the second with a smaller grid, in FP32 + FP64 band, in FP64,
and in the AP mode whose overflowed cells go through the adaptive list."""
import ctypes

import numpy as np
import pytest
import torch

from quijibo_amd import _lib, api, workloads as W
from quijibo_amd._lib import FP32_BAND, FP64, MAP_ADAPTIVE, MAP_BLACK_HOLE, MAP_SKIP_ORIGIN, check

pytestmark = pytest.mark.gpu


def _noalloc(grid, r, p, prec, band, out):
    return _lib.load().qj_field_eval_noalloc(ctypes.byref(grid), r.handle, ctypes.byref(p), int(prec), float(band),
                                             ctypes.c_void_p(out.data_ptr()),
                                             ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def _both(w, res, prec, flags=MAP_BLACK_HOLE, escape=None, max_iter=None):
    g = api.make_grid(res, w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers), api.POLYNOMIAL_4D(w.bottom.roots, w.bottom.powers))
    p = api.make_params(w.max_iter if max_iter is None else max_iter, w.escape if escape is None else escape,
                        w.exp_scaling, w.slice, flags)
    n = res[0] * res[1] * res[2]
    ref = torch.full((n,), 7.25, dtype=torch.float64, device="cuda")
    api.field_eval(g, r, p, prec, 0.05, out=ref)
    out = torch.full((n,), 7.25, dtype=torch.float64, device="cuda")
    print("noalloc %s prec %d flags %d" % (res, prec, flags), flush=True)
    check(_noalloc(g, r, p, prec, 0.05, out))
    torch.cuda.synchronize()
    a, b = out.cpu().numpy(), ref.cpu().numpy()
    assert np.array_equal(a, b, equal_nan=True), res
    return a


@pytest.mark.parametrize("prec", [FP32_BAND, FP64])
def test_noalloc_matches_field_eval_on_two_calls(gpu, prec):
    check(_lib.load().qj_init(0))
    w = W.config2(8)
    # 256 x 256 x 80 = 5.2M cells: two z-slabs of the 2^22-cell scratch
    f = _both(w, (256, 256, 80), prec)
    assert np.isfinite(f).mean() > 0.9
    _both(w, (48, 40, 36), prec)


def test_noalloc_adaptive_list(gpu):
    """AP semantics (escape 20, origin skipped, overflowed cells re-passed in
    extended range) with roots whose products overflow FP64."""
    top = W.RootSet(np.array([[1.9, 0.3, 0.1, 0.0], [-2.1, 0.4, -0.2, 0.0], [0.2, -0.3, 0.1, 0.05]]),
                    np.array([200.0, 200.0, 1.0]))
    bot = W.RootSet(np.array([[1.902, 0.31, 0.1, 0.0], [-2.1, 0.41, -0.19, 0.0], [0.3, 0.2, -0.1, 0.0]]),
                    np.array([200.0, 200.0, 1.0]))
    w = W.Workload("ap", (40, 36, 33), top, bot, max_iter=1, escape=20.0, exp_scaling=0.5)
    _both(w, (40, 36, 33), FP64, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)
    _both(w, (20, 18, 17), FP64, MAP_SKIP_ORIGIN | MAP_ADAPTIVE)


def test_noalloc_rejects_oversized_slices(gpu):
    w = W.config1(8)
    g = api.make_grid((4096, 1025, 1), w.center, w.lengths)
    r = api.Rational(api.POLYNOMIAL_4D(w.top.roots, w.top.powers))
    p = api.make_params(3, 2000.0, 1.0, 0.0, MAP_BLACK_HOLE)
    out = torch.zeros(8, dtype=torch.float64, device="cuda")
    assert _noalloc(g, r, p, FP32_BAND, 0.05, out) == 4  # QJ_ERR_CAPACITY, nothing launched
    assert b"exceeds the maximum" in _lib.load().qj_last_error()
