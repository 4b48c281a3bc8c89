"""The Python layer rejects tensors the C ABI would receive as bad device
pointers: host tensors, wrong dtypes, non-contiguous views and wrong sizes.
The checks run before any device call, so these tests need no GPU."""
import pytest
import torch

from quijibo_amd import api


def _grid(res=(4, 5, 6)):
    return api.make_grid(res)


def test_field_eval_rejects_host_and_missized_out():
    with pytest.raises(ValueError, match="CUDA"):
        api.field_eval(_grid(), None, None, out=torch.zeros(120, dtype=torch.float64))


def test_points_eval_rejects_host_points():
    with pytest.raises(ValueError, match="CUDA"):
        api.points_eval(torch.zeros((10, 3), dtype=torch.float64), None, None)
    with pytest.raises(ValueError, match=r"\(n, 3\)"):
        api.points_eval(torch.zeros((10, 4), dtype=torch.float64), None, None)


def test_marching_cubes_rejects_host_field():
    with pytest.raises(ValueError, match="CUDA"):
        api.marching_cubes(torch.zeros(120, dtype=torch.float64), _grid())


def test_ap_gradients_reject_host_field():
    with pytest.raises(ValueError, match="CUDA"):
        api.ap_score_gradients(_grid(), None, None, _grid(), torch.zeros(120, dtype=torch.float64),
                               field=torch.zeros(120, dtype=torch.float64))


def test_points_adjoint_rejects_host_points():
    with pytest.raises(ValueError, match="CUDA"):
        api.points_adjoint(torch.zeros((10, 3), dtype=torch.float64), None, None, torch.zeros(10))


def test_device_tensor_checks_dtype_contiguity_and_size():
    # the checker itself, on a CPU tensor standing in for the device check's later stages
    t = torch.zeros(12, dtype=torch.float64)
    with pytest.raises(ValueError, match="CUDA"):
        api._device_tensor(t, "x", 12)
    with pytest.raises(ValueError, match="torch tensor"):
        api._device_tensor(t.numpy(), "x")


def test_scoring_fields_reject_host_tensors():
    with pytest.raises(ValueError, match="CUDA"):
        api.inverted_distance(torch.zeros(120, dtype=torch.float64), _grid())
    with pytest.raises(ValueError, match="CUDA"):
        api.weight_scores(torch.zeros(120, dtype=torch.float64), _grid(), torch.zeros(120, dtype=torch.float64),
                          _grid(), 10.0)
    with pytest.raises(ValueError, match="CUDA"):
        api.field_min_max(torch.zeros(120, dtype=torch.float64))
