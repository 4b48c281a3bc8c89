"""The C-ABI library loads without a GPU and exports every symbol that
include/qj_c_api.h declares (no compute calls here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "qj_c_api.h")


def declared():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s+(qj_[a-z_0-9]+)\s*\(", txt, flags=re.M)))


def test_header_declares_the_api():
    names = declared()
    for must in ["qj_field_eval", "qj_points_eval", "qj_mc_run", "qj_mc_copy", "qj_ap_score_gradients",
                 "qj_points_adjoint", "qj_rational_create", "qj_workspace_create", "qj_server_create",
                 "qj_server_submit", "qj_server_poll"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from quijibo_amd import _lib
    L = _lib.load()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing
    assert {n for n, _, _ in _lib.SIGNATURES} == set(declared())


def test_version_and_status_strings():
    from quijibo_amd import _lib
    L = _lib.load()
    assert b"gfx950" in L.qj_version()
    assert L.qj_status_string(0) == b"ok"
    assert L.qj_status_string(4) == b"capacity exceeded"


def test_invalid_arguments_are_status_codes():
    from quijibo_amd import _lib
    L = _lib.load()
    out = ctypes.c_void_p()
    assert L.qj_rational_create(None, None, ctypes.byref(out)) == 1  # QJ_ERR_INVALID, no device touched
    assert b"null" in L.qj_last_error()
