"""ctypes bindings to libquijibo.so (the C ABI in include/qj_c_api.h).

The library is the only compute path: if it is missing or fails to load,
importing this module raises, so nothing silently falls back to the CPU.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libquijibo.so")

QJ_OK = 0
FP64 = 0
FP32_BAND = 1
FP64_EXACT = 2
MAP_BLACK_HOLE = 1
MAP_SKIP_ORIGIN = 2
MAP_ADAPTIVE = 4
MAP_LAGRANGE = 8
MAP_EXTENDED_RANGE = 16

c_double_p = ctypes.POINTER(ctypes.c_double)


class QJError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("quijibo status %d: %s" % (status, msg))
        self.status = status


class Grid(ctypes.Structure):
    _fields_ = [("res", ctypes.c_int32 * 3), ("center", ctypes.c_double * 3),
                ("lengths", ctypes.c_double * 3), ("rotation", ctypes.c_double * 4)]


class Poly(ctypes.Structure):
    _fields_ = [("n_roots", ctypes.c_int32), ("roots", c_double_p), ("powers", c_double_p),
                ("power_scalar", ctypes.c_double)]


class MapParams(ctypes.Structure):
    _fields_ = [("max_iter", ctypes.c_int32), ("escape", ctypes.c_double), ("exp_scaling", ctypes.c_double),
                ("slice", ctypes.c_double), ("flags", ctypes.c_uint32)]


class EvalStats(ctypes.Structure):
    _fields_ = [("voxels", ctypes.c_int64), ("iterations", ctypes.c_int64), ("band_voxels", ctypes.c_int64),
                ("band_iterations", ctypes.c_int64), ("adaptive_voxels", ctypes.c_int64),
                ("nonfinite", ctypes.c_int64), ("checked", ctypes.c_int64), ("check_sign_flips", ctypes.c_int64),
                ("check_class_mismatch", ctypes.c_int64), ("check_max_abs_dF", ctypes.c_double),
                ("check_max_rel_dF", ctypes.c_double)]


class APGrads(ctypes.Structure):
    _fields_ = [("score", ctypes.c_double), ("grad_exp", ctypes.c_double), ("grad_top_scalar", ctypes.c_double),
                ("grad_bottom_scalar", ctypes.c_double), ("adaptive_cells", ctypes.c_int64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget_bytes", ctypes.c_int64), ("chunk_voxels", ctypes.c_int32),
                ("chunks_per_launch", ctypes.c_int32), ("max_roots", ctypes.c_int32),
                ("inflight_launches", ctypes.c_int32), ("precision", ctypes.c_int32), ("band", ctypes.c_double),
                ("device", ctypes.c_int32), ("launch_work", ctypes.c_double),
                ("max_pending_work", ctypes.c_double), ("aging_s", ctypes.c_double),
                ("check_samples", ctypes.c_int32), ("escalate_abs_dF", ctypes.c_double)]


class Request(ctypes.Structure):
    _fields_ = [("grid", Grid), ("top", Poly), ("bottom", Poly), ("params", MapParams),
                ("keep_field", ctypes.c_int32), ("arrival_time", ctypes.c_double)]


class Completion(ctypes.Structure):
    _fields_ = [("id", ctypes.c_uint64), ("status", ctypes.c_int32), ("arrival_time", ctypes.c_double),
                ("admit_time", ctypes.c_double), ("done_time", ctypes.c_double), ("d_field", ctypes.c_void_p),
                ("voxels", ctypes.c_int64), ("checked", ctypes.c_int64), ("check_sign_flips", ctypes.c_int64),
                ("check_class_mismatch", ctypes.c_int64), ("check_max_abs_dF", ctypes.c_double),
                ("check_max_rel_dF", ctypes.c_double), ("escalated", ctypes.c_int32)]


# every symbol declared in include/qj_c_api.h: (name, restype, argtypes)
_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = [
    ("qj_version", ctypes.c_char_p, []),
    ("qj_last_error", ctypes.c_char_p, []),
    ("qj_status_string", ctypes.c_char_p, [ctypes.c_int32]),
    ("qj_init", ctypes.c_int32, [ctypes.c_int32]),
    ("qj_rational_create", ctypes.c_int32, [ctypes.POINTER(Poly), ctypes.POINTER(Poly), _PP]),
    ("qj_rational_update", ctypes.c_int32, [_P, ctypes.POINTER(Poly), ctypes.POINTER(Poly), _P]),
    ("qj_rational_counts", ctypes.c_int32, [_P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]),
    ("qj_rational_destroy", ctypes.c_int32, [_P]),
    ("qj_workspace_create", ctypes.c_int32, [_PP]),
    ("qj_workspace_destroy", ctypes.c_int32, [_P]),
    ("qj_workspace_stats", ctypes.c_int32, [_P, _P, ctypes.POINTER(EvalStats)]),
    ("qj_workspace_reset_stats", ctypes.c_int32, [_P, _P]),
    ("qj_workspace_check", ctypes.c_int32, [_P, ctypes.c_int64]),
    ("qj_workspace_timing", ctypes.c_int32, [_P, ctypes.c_int32]),
    ("qj_workspace_kernel_ms", ctypes.c_int32, [_P, c_double_p, c_double_p, ctypes.POINTER(ctypes.c_int64)]),
    ("qj_field_eval", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams), ctypes.c_int32,
                                       ctypes.c_double, _P, _P]),
    ("qj_field_eval_noalloc", ctypes.c_int32, [ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams), ctypes.c_int32,
                                               ctypes.c_double, _P, _P]),
    ("qj_points_eval", ctypes.c_int32, [_P, _P, ctypes.c_int64, _P, ctypes.POINTER(MapParams), ctypes.c_int32,
                                        ctypes.c_double, _P, _P]),
    ("qj_mc_run", ctypes.c_int32, [_P, _P, ctypes.POINTER(Grid), ctypes.c_double, ctypes.POINTER(ctypes.c_int64),
                                   ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_mc_run_tiles", ctypes.c_int32, [_P, ctypes.c_int32, _P, ctypes.POINTER(Grid), ctypes.c_double,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_mc_run_tiles_into", ctypes.c_int32, [_P, ctypes.c_int32, _P, ctypes.POINTER(Grid), ctypes.c_double, _P,
                                              ctypes.c_int64, _P, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                              ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_nlmc_run", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams), ctypes.c_double,
                                     ctypes.c_int32, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_int64),
                                     ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_nlmc_field", ctypes.c_int32, [_P, _P, ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams), ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_nlmc_cache", ctypes.c_int32, [_P, ctypes.POINTER(ctypes.c_int64), _P, ctypes.POINTER(ctypes.c_int64), _P,
                                       _P]),
    ("qj_nlmc_resume", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams), ctypes.c_double,
                                        ctypes.c_int64, _P, ctypes.c_int64, _P, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), _P]),
    ("qj_mc_copy", ctypes.c_int32, [_P, _P, _P, _P]),
    ("qj_mc_device_ptrs", ctypes.c_int32, [_P, _PP, _PP]),
    ("qj_field_clamp_infs_to_mean", ctypes.c_int32, [_P, _P, ctypes.c_int64, c_double_p, _P]),
    ("qj_field_clamp_nans_to_mean", ctypes.c_int32, [_P, _P, ctypes.c_int64, c_double_p, _P]),
    ("qj_field_set_borders", ctypes.c_int32, [ctypes.POINTER(Grid), ctypes.c_double, _P, _P]),
    ("qj_field_min_max", ctypes.c_int32, [_P, _P, ctypes.c_int64, c_double_p, c_double_p, _P]),
    ("qj_field_inverted_distance", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.c_int32, _P, c_double_p,
                                                    _P]),
    ("qj_weight_scores", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.POINTER(Grid), _P, ctypes.c_double,
                                          c_double_p, ctypes.c_int32, c_double_p, c_double_p, _P]),
    ("qj_ap_score_gradients", ctypes.c_int32, [_P, ctypes.POINTER(Grid), _P, ctypes.POINTER(MapParams),
                                               ctypes.POINTER(Grid), _P, ctypes.c_double, ctypes.c_double, _P, _P,
                                               _P, ctypes.POINTER(APGrads), _P]),
    ("qj_points_adjoint", ctypes.c_int32, [_P, _P, ctypes.c_int64, _P, ctypes.POINTER(MapParams), ctypes.c_int32,
                                           _P, _P, _P, _P]),
    ("qj_server_create", ctypes.c_int32, [ctypes.POINTER(ServerConfig), _PP]),
    ("qj_server_now", ctypes.c_double, [_P]),
    ("qj_server_submit", ctypes.c_int32, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_uint64)]),
    ("qj_server_poll", ctypes.c_int32, [_P, ctypes.POINTER(Completion), ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int32), ctypes.c_double]),
    ("qj_server_fetch", ctypes.c_int32, [_P, ctypes.POINTER(Completion), _P, _P]),
    ("qj_server_release", ctypes.c_int32, [_P, ctypes.c_uint64]),
    ("qj_server_destroy", ctypes.c_int32, [_P]),
]

_lib = None


def load():
    """Load libquijibo.so; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError("libquijibo.so is not built (%s); run __graft_entry__.build() or "
                          "python -m quijibo_amd.build" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status):
    if status != QJ_OK:
        L = load()
        raise QJError(status, (L.qj_last_error() or b"").decode() or L.qj_status_string(status).decode())
    return status
