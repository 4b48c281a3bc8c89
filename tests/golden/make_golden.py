"""Writes tests/golden/oracle_vectors.npz: small seeded inputs with the oracle's outputs,
frozen as regression vectors, plus the one known-answer case the reference's own code
holds (POLYNOMIAL_4D::rationalTest, POLYNOMIAL_4D.cpp:520-577: (z^2 - 1) / z^2 at
z = 1.23 + 4.56i equals 1 - 1/z^2).

The reference ships no fields, meshes or .o3d files, so apart from rationalTest these
vectors come from the oracle; tests check the oracle still reproduces them bit for bit
and the GPU path matches them to the FP64 tolerance. Run from the repository root:
    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

import oracle as O  # noqa: E402
from quijibo_amd import workloads as W  # noqa: E402


def main():
    out = {}
    for name, w, res in (("c1", W.config1(16), (16, 16, 16)), ("c2", W.config2(12), (12, 11, 10))):
        g = O.make_grid(res, w.center, w.lengths)
        top, bot = O.poly(w.top.roots, w.top.powers), O.poly(w.bottom.roots, w.bottom.powers)
        f = O.field_map(g, top, bot, O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice, O.BLACK_HOLE))
        out[name + "_res"] = np.array(res)
        out[name + "_field"] = f
        v, t = O.marching_cubes(g, f)
        out[name + "_mc_vertices"] = v
        out[name + "_mc_triangles"] = t
        lf = O.field_map(g, top, bot, O.make_params(w.max_iter, 2000.0, w.exp_scaling, w.slice,
                                                    O.BLACK_HOLE | O.LAGRANGE))
        out[name + "_lagrange_field"] = lf
    # rationalTest: top roots +1 and -1 (z^2 - 1), bottom double root 0 (z^2), one iteration
    z = complex(1.23, 4.56)
    out["rational_point"] = np.array([1.23, 4.56, 0.0])
    out["rational_log_abs_closed_form"] = np.array([np.log(abs(1 - 1 / z ** 2))])
    top = O.poly(np.array([[1.0, 0, 0, 0], [-1.0, 0, 0, 0]]))
    bot = O.poly(np.array([[0.0, 0, 0, 0]]), np.array([2.0]))
    out["rational_oracle"] = O.points_eval(out["rational_point"][None, :], top, bot,
                                           O.make_params(1, 2000.0, 1.0, 0.0, O.BLACK_HOLE))
    np.savez(os.path.join(HERE, "oracle_vectors.npz"), **out)
    print("wrote", sorted(out))


if __name__ == "__main__":
    main()
