"""Pins the oracle's scoring fields of the shape optimizer:

* invertedDistance (FIELD_3D.cpp:5241-5297): the sweep numbering is 1 + the
  6-neighbour (taxicab) distance to the cells next to a sign change, so scipy's
  chamfer distance transform with the taxicab metric, from a shell computed
  here with numpy over the 26 neighbours, reproduces it bit for bit;
* invertedDistanceNormalized (:5316-5352): inside and outside each sum to 0.5
  in absolute value and stay proportional to the unnormalized field;
* weightScore / inverseDistanceScore (OPTIMIZE_3D.cpp:4502-4563) with the
  weight on the fractal's own grid (the lookup at a cell centre returns the
  cell's value): the direct sum, its sign and threshold symmetries, and the
  maximum score as the sum of |w|;
* fieldMin / fieldMax (FIELD_3D.cpp:2960-3032): numpy's nan-aware extrema,
  NaN when cell 0 is NaN.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O


def _sphere(res, lengths=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), r=0.3, c=(0.05, -0.02, 0.03)):
    ax = [center[i] - lengths[i] / 2 + (np.arange(res[i]) + 0.5) * lengths[i] / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return (np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2) - r).reshape(-1)


def _shell(d, res):
    """Cells with a 26-neighbour of the other sign (cells at 0 have none)."""
    a = d.reshape(res[2], res[1], res[0])
    out = np.zeros(a.shape, bool)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                if dx == dy == dz == 0:
                    continue
                sl_dst, sl_src = [], []
                for o, n in ((dz, a.shape[0]), (dy, a.shape[1]), (dx, a.shape[2])):
                    sl_dst.append(slice(max(0, -o), n - max(0, o)))
                    sl_src.append(slice(max(0, o), n - max(0, -o)))
                out[tuple(sl_dst)] |= (a[tuple(sl_dst)] * a[tuple(sl_src)]) < 0
    return out


def _reference_inverted(d, res):
    layer = 1.0 + ndimage.distance_transform_cdt(~_shell(d, res), metric="taxicab").astype(np.float64)
    v = 1.0 / layer.reshape(-1)
    return np.where(d < 0, -v, 0.25 * v)


@pytest.mark.parametrize("res", [(12, 10, 9), (7, 13, 5), (16, 16, 16)])
def test_inverted_distance_is_one_over_one_plus_taxicab_distance(res):
    d = _sphere(res)
    g = O.make_grid(res, (0, 0, 0), (1, 1, 1))
    got, _ = O.inverted_distance(g, d)
    assert np.array_equal(got, _reference_inverted(d, res))


def test_inverted_distance_planted_mistake_fails():
    """A 6-neighbour shell (instead of 26) changes the result, so the check
    above would catch it."""
    res = (12, 10, 9)
    d = _sphere(res)
    a = d.reshape(res[2], res[1], res[0])
    six = np.zeros(a.shape, bool)
    for ax in range(3):
        for o in (-1, 1):
            b = np.roll(a, o, axis=ax)
            m = a * b < 0
            idx = [slice(None)] * 3
            idx[ax] = 0 if o == 1 else -1
            m[tuple(idx)] = False
            six |= m
    layer = 1.0 + ndimage.distance_transform_cdt(~six, metric="taxicab").astype(np.float64)
    wrong = np.where(d < 0, -1.0, 0.25) / layer.reshape(-1)
    got, _ = O.inverted_distance(O.make_grid(res, (0, 0, 0), (1, 1, 1)), d)
    assert not np.array_equal(got, wrong)


def test_inverted_distance_normalized_sums():
    res = (14, 12, 10)
    d = _sphere(res)
    g = O.make_grid(res, (0, 0, 0), (1, 1, 1))
    raw, _ = O.inverted_distance(g, d)
    nrm, a = O.inverted_distance(g, d, normalized=True)
    inside = d < 0
    # serial sums against numpy's pairwise ones: within n u of the total
    tol = nrm.size * 2.0 ** -53
    assert abs(nrm[inside].sum() + 0.5) < tol and abs(nrm[~inside].sum() - 0.5) < tol
    assert abs(a - np.abs(nrm).sum()) < tol and abs(a - 1.0) < tol
    for m in (inside, ~inside):
        ratio = nrm[m] / raw[m]
        assert np.allclose(ratio, ratio[0], rtol=1e-14, atol=0)


def test_inverted_distance_without_a_surface_is_an_error():
    res = (6, 5, 4)
    with pytest.raises(ValueError):
        O.inverted_distance(O.make_grid(res, (0, 0, 0), (1, 1, 1)), np.ones(120))


def test_weight_score_on_the_fractal_grid():
    rng = np.random.default_rng(11)
    res = (9, 8, 7)
    g = O.make_grid(res, (0.1, -0.2, 0.3), (2.0, 1.5, 1.2))
    f = rng.normal(size=res[0] * res[1] * res[2])
    w = rng.normal(size=f.size)
    for thr in (0.0, 0.37, -1.2):
        s, m = O.weight_score(g, f, g, w, 10.0, thr)
        assert abs(s - np.sum(-np.tanh(10.0 * (f - thr)) * w)) <= 1e-12 * np.abs(w).sum()
        assert abs(m - np.abs(w).sum()) <= 1e-13 * m
    # antisymmetry about the threshold and linearity in the weight
    s1, _ = O.weight_score(g, f, g, w, 7.0, 0.2)
    s2, _ = O.weight_score(g, 0.4 - f, g, w, 7.0, 0.2)
    assert abs(s1 + s2) <= 1e-12 * np.abs(w).sum()
    w2 = rng.normal(size=f.size)
    sa, _ = O.weight_score(g, f, g, w + w2, 5.0, 0.1)
    sb, _ = O.weight_score(g, f, g, w, 5.0, 0.1)
    sc, _ = O.weight_score(g, f, g, w2, 5.0, 0.1)
    assert abs(sa - sb - sc) <= 1e-12 * (np.abs(w).sum() + np.abs(w2).sum())
    # a transposed weight (z fastest instead of x) scores differently
    wt = w.reshape(res[2], res[1], res[0]).transpose(2, 1, 0).reshape(-1)
    assert abs(O.weight_score(g, f, g, wt, 10.0, 0.0)[0] - O.weight_score(g, f, g, w, 10.0, 0.0)[0]) > 1e-6


def test_field_min_max_semantics():
    rng = np.random.default_rng(5)
    f = rng.normal(size=500)
    f[[17, 230]] = np.nan
    mn, mx = O.field_min_max(f)
    assert mn == np.nanmin(f) and mx == np.nanmax(f)
    g = f.copy()
    g[0] = np.nan
    assert all(np.isnan(v) for v in O.field_min_max(g))
    z = np.array([0.0, 1.0, -0.0, -1.0, 2.0, -1.0])
    assert O.field_min_max(z) == (-1.0, 2.0)
    z = np.array([0.0, -0.0, 3.0])
    mn, _ = O.field_min_max(z)
    assert mn == 0.0 and np.copysign(1.0, mn) == 1.0  # ties keep the first cell
