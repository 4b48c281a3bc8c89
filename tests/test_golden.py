"""The committed vectors (tests/golden/oracle_vectors.npz, written by
tests/golden/make_golden.py): the oracle still reproduces them bit for bit, and the
rationalTest known answer (POLYNOMIAL_4D.cpp:520-577) holds."""
import os

import numpy as np
import pytest

import oracle as O
from quijibo_amd import workloads as W

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "oracle_vectors.npz"))


def test_rational_test_known_answer():
    assert abs(GOLD["rational_oracle"][0] - GOLD["rational_log_abs_closed_form"][0]) < 1e-14
    z = complex(1.23, 4.56)
    assert abs(GOLD["rational_log_abs_closed_form"][0] - np.log(abs((z * z - 1) / (z * z)))) < 1e-15


@pytest.mark.parametrize("name", ["c1", "c2"])
def test_oracle_reproduces_the_vectors(name):
    w = W.config1(16) if name == "c1" else W.config2(12)
    res = tuple(int(v) for v in GOLD[name + "_res"])
    g = O.make_grid(res, w.center, w.lengths)
    top, bot = O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers)
    f = O.field_map(g, top, bot, O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
    assert np.array_equal(f, GOLD[name + "_field"], equal_nan=True)
    v, t = O.marching_cubes(g, f)
    assert np.array_equal(v, GOLD[name + "_mc_vertices"], equal_nan=True)
    assert np.array_equal(t, GOLD[name + "_mc_triangles"])
    lf = O.field_map(g, top, bot, O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice,
                                                O.BLACK_HOLE | O.LAGRANGE))
    assert np.array_equal(lf, GOLD[name + "_lagrange_field"], equal_nan=True)
