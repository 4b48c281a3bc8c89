"""quijibo_amd: MI355X-native (gfx950 HIP) rational-quaternion Julia field path.

Field evaluation, marching cubes, gradients and the request loop run in
libquijibo.so (HIP kernels behind the C ABI in include/qj_c_api.h).
"""
from ._lib import FP32_BAND, FP64, MAP_ADAPTIVE, MAP_BLACK_HOLE, MAP_SKIP_ORIGIN, QJError, load  # noqa: F401

__all__ = ["FP64", "FP32_BAND", "MAP_BLACK_HOLE", "MAP_SKIP_ORIGIN", "MAP_ADAPTIVE", "QJError", "load"]
