"""The atan polynomial in qj_fp32.h::angle_upper: its coefficients are the ones
quijibo_amd/csrc/fit_atan.py fits, and the FP32 evaluation stays within 1.7
ulp of atan over [0, 1] and gives atan2 over the upper half plane to a few
float32 ulps of pi."""
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "quijibo_amd", "csrc")
sys.path.insert(0, CSRC)
import fit_atan  # noqa: E402


def _kernel_coefficients():
    src = open(os.path.join(CSRC, "qj_fp32.h")).read()
    body = src[src.index("float angle_upper("):]
    body = body[:body.index("float r = t * p;")]
    first = float(re.search(r"float p = ([-0-9.e]+)f;", body).group(1))
    rest = [float(x) for x in re.findall(r"__builtin_fmaf\(p, s, ([-0-9.e]+)f\)", body)]
    return np.array([first] + rest, dtype=np.float32)[::-1]  # lowest degree first


def test_kernel_coefficients_are_the_fit():
    assert np.array_equal(_kernel_coefficients(), fit_atan.coefficients())


def test_fp32_error_bound():
    assert fit_atan.max_ulp(_kernel_coefficients()) < 1.7


def test_upper_half_plane_angle():
    c = _kernel_coefficients()
    rng = np.random.default_rng(3)
    y = np.abs(rng.normal(size=200000)).astype(np.float32)
    w = rng.normal(size=200000).astype(np.float32)
    aw = np.abs(w)
    steep = y > aw
    t = (np.where(steep, aw, y) / np.where(steep, y, aw)).astype(np.float32)
    r = fit_atan.horner32(c, t)
    r = np.where(steep, np.float32(np.pi / 2) - r, r).astype(np.float32)
    r = np.where(w < 0, np.float32(np.pi) - r, r).astype(np.float32)
    ref = np.arctan2(y.astype(np.float64), w.astype(np.float64))
    assert np.max(np.abs(r - ref)) < 4 * np.spacing(np.float32(np.pi))
