"""The shape optimizer's scoring fields on the GPU against the oracle:
qj_field_inverted_distance (FIELD_3D::invertedDistance / Normalized),
qj_weight_scores (weightScore / inverseDistanceScore / the maximum score) and
qj_field_min_max (fieldMin / fieldMax).

Tolerances: the layer numbering and the plain inverted field bit for bit (the
values are 1/n, negated or quartered: exact operations on the same integers);
normalized fields and scores within n u relative (u = 2^-53; device sums are
double-double, the oracle's serial); min / max exact."""
import math

import numpy as np
import pytest
import torch

import oracle as O
from quijibo_amd import api
from quijibo_amd._lib import QJError

pytestmark = pytest.mark.gpu


def _ellipsoid(res, center, lengths):
    ax = [center[i] - lengths[i] / 2 + (np.arange(res[i]) + 0.5) * lengths[i] / res[i] for i in range(3)]
    Z, Y, X = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    return (np.sqrt((X / 1.0) ** 2 + (Y / 0.7) ** 2 + (Z / 0.85) ** 2) - 0.55).reshape(-1)


@pytest.mark.parametrize("res", [(20, 18, 16), (33, 7, 29), (64, 64, 64)])
def test_inverted_distance_against_the_oracle(gpu, res):
    center, lengths = (0.1, -0.05, 0.02), (2.0, 1.8, 1.9)
    d = _ellipsoid(res, center, lengths)
    g = api.make_grid(res, center, lengths)
    og = O.make_grid(res, center, lengths)
    got, _ = api.inverted_distance(torch.from_numpy(d).cuda(), g)
    ref, _ = O.inverted_distance(og, d)
    assert np.array_equal(got.cpu().numpy(), ref)
    gn, a = api.inverted_distance(torch.from_numpy(d).cuda(), g, normalized=True)
    rn, ra = O.inverted_distance(og, d, normalized=True)
    tol = d.size * 2.0 ** -52
    gn = gn.cpu().numpy()
    print("normalized: max rel %.2e, abs sums %.17g / %.17g" % (np.max(np.abs(gn - rn) / np.abs(rn)), a, ra),
          flush=True)
    assert np.max(np.abs(gn - rn) / np.abs(rn)) <= tol and abs(a - ra) <= tol
    inside = d < 0
    assert abs(gn[inside].sum() + 0.5) <= tol and abs(gn[~inside].sum() - 0.5) <= tol


def test_inverted_distance_without_a_surface_is_an_error(gpu):
    res = (8, 7, 6)
    g = api.make_grid(res, (0, 0, 0), (1, 1, 1))
    with pytest.raises(QJError) as e:
        api.inverted_distance(torch.ones(8 * 7 * 6, dtype=torch.float64, device="cuda"), g)
    assert e.value.status == 5


def test_weight_scores_against_the_oracle(gpu):
    rng = np.random.default_rng(9)
    res, wres = (25, 25, 25), (20, 18, 22)
    center, lengths = (0.0, 0.1, -0.1), (4.0, 4.0, 4.0)
    g, og = api.make_grid(res, center, lengths), O.make_grid(res, center, lengths)
    wg, owg = api.make_grid(wres, (0.05, 0.0, 0.0), (3.6, 3.4, 3.8)), O.make_grid(wres, (0.05, 0.0, 0.0),
                                                                                  (3.6, 3.4, 3.8))
    f = rng.normal(scale=0.7, size=int(np.prod(res)))
    w, _ = O.inverted_distance(owg, _ellipsoid(wres, (0.05, 0.0, 0.0), (3.6, 3.4, 3.8)), normalized=True)
    thr = np.linspace(-1.3, 1.1, 101)
    sc, mx = api.weight_scores(torch.from_numpy(f).cuda(), g, torch.from_numpy(w).cuda(), wg, 10.0, thr)
    tol = f.size * 2.0 ** -52
    worst = 0.0
    for k, t in enumerate(thr):
        rs, rm = O.weight_score(og, f, owg, w, 10.0, t)
        worst = max(worst, abs(sc[k] - rs) / rm)
        assert abs(mx - rm) <= tol * rm
    print("101 thresholds: max |score - oracle| / max score %.2e" % worst, flush=True)
    assert worst <= tol
    # no threshold: the maximum score alone
    sc0, mx0 = api.weight_scores(torch.from_numpy(f).cuda(), g, torch.from_numpy(w).cuda(), wg, 10.0, [])
    assert sc0.size == 0 and mx0 == mx


def test_field_min_max_semantics(gpu):
    rng = np.random.default_rng(2)
    f = rng.normal(size=300001)
    f[[5, 77777, 300000]] = np.nan
    f[123456] = -np.inf
    got = api.field_min_max(torch.from_numpy(f).cuda())
    assert got == O.field_min_max(f)
    g = f.copy()
    g[0] = np.nan
    assert all(math.isnan(v) for v in api.field_min_max(torch.from_numpy(g).cuda()))
    z = np.zeros(100000)
    z[1::2] = -0.0
    mn, mx = api.field_min_max(torch.from_numpy(z).cuda())
    assert mn == 0.0 and math.copysign(1.0, mn) == 1.0 and math.copysign(1.0, mx) == 1.0
    z[0] = -0.0
    mn, mx = api.field_min_max(torch.from_numpy(z).cuda())
    assert math.copysign(1.0, mn) == -1.0 and math.copysign(1.0, mx) == -1.0
