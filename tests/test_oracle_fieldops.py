"""Pins the oracle's FIELD_3D clean-up passes (oracle/qo.c clamp_to_mean,
qo_set_borders; FIELD_3D.cpp:5182-5234, 5357-5382): exact means on fields
whose serial sums are exact, the serial sum's non-finite outcomes, the
serial sum's error bound against a correctly rounded sum (math.fsum), and the
six faces against numpy slicing."""
import math

import numpy as np
import pytest

import oracle as O

U = 2.0 ** -53


def test_clamp_infs_exact_mean_and_untouched_cells():
    rng = np.random.default_rng(1)
    f = rng.integers(-1000, 1000, size=5000).astype(np.float64)
    inf_at = rng.choice(5000, 300, replace=False)
    f[inf_at[:150]] = np.inf
    f[inf_at[150:]] = -np.inf
    out, mean = O.clamp_infs_to_mean(f)
    keep = np.isfinite(f)
    assert mean == np.sum(f[keep]) / keep.sum()  # integer sums are exact in any order
    assert np.array_equal(out[keep], f[keep])
    assert np.all(out[~keep] == mean)


def test_clamp_nans_exact_mean():
    f = np.arange(100, dtype=np.float64)
    f[[3, 50, 99]] = np.nan
    out, mean = O.clamp_nans_to_mean(f)
    keep = ~np.isnan(f)
    assert mean == np.sum(f[keep]) / keep.sum()
    assert np.all(out[[3, 50, 99]] == mean) and not np.isnan(out).any()


def test_serial_sum_quirks():
    # a NaN cell is "not infinite", so clampInfsToMean's mean becomes NaN
    out, mean = O.clamp_infs_to_mean([1.0, np.nan, np.inf, 2.0])
    assert math.isnan(mean) and np.isnan(out[2]) and out[0] == 1.0
    # an infinite cell is "not NaN", so clampNansToMean's mean is infinite
    out, mean = O.clamp_nans_to_mean([1.0, np.nan, np.inf, 2.0])
    assert mean == np.inf and out[1] == np.inf
    out, mean = O.clamp_nans_to_mean([1.0, np.nan, np.inf, -np.inf])
    assert math.isnan(mean)
    # nothing summed: 0 / 0
    out, mean = O.clamp_infs_to_mean([np.inf, -np.inf])
    assert math.isnan(mean) and np.all(np.isnan(out))


def test_serial_mean_within_its_error_bound():
    rng = np.random.default_rng(2)
    f = rng.normal(size=200000) * np.exp(rng.uniform(-20, 20, size=200000))
    f[::997] = -np.inf
    keep = np.isfinite(f)
    _, mean = O.clamp_infs_to_mean(f)
    exact = math.fsum(f[keep]) / keep.sum()
    bound = (keep.sum() - 1) * U * np.sum(np.abs(f[keep])) / keep.sum()
    assert abs(mean - exact) <= bound
    # a planted mistake: the mean over every cell (count including the infinite ones)
    assert abs(math.fsum(f[keep]) / f.size - exact) > bound


@pytest.mark.parametrize("res", [(5, 4, 3), (1, 6, 7), (2, 2, 2), (9, 1, 1)])
def test_set_borders_faces(res):
    g = O.make_grid(res)
    f = np.arange(np.prod(res), dtype=np.float64)
    out = O.set_borders(g, f, -7.5).reshape(res[2], res[1], res[0])
    ref = f.reshape(res[2], res[1], res[0]).copy()
    ref[0, :, :] = ref[-1, :, :] = -7.5
    ref[:, 0, :] = ref[:, -1, :] = -7.5
    ref[:, :, 0] = ref[:, :, -1] = -7.5
    assert np.array_equal(out, ref)
