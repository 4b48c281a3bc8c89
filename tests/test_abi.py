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


def test_ctypes_structs_match_the_header(tmp_path):
    """Every struct the Python layer mirrors has the C header's size and field
    offsets (gcc on the header prints them), so a struct the library writes
    never runs past the buffer Python handed it."""
    import shutil
    import subprocess
    from quijibo_amd import _lib
    if not shutil.which("gcc"):
        pytest.skip("gcc not available")
    pairs = [(_lib.Grid, "qj_grid"), (_lib.Poly, "qj_poly"), (_lib.MapParams, "qj_map_params"),
             (_lib.EvalStats, "qj_eval_stats"), (_lib.APGrads, "qj_ap_grads"),
             (_lib.ServerConfig, "qj_server_config"), (_lib.Request, "qj_request"),
             (_lib.Completion, "qj_completion")]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "qj_c_api.h"', "int main(void) {"]
    for cls, cname in pairs:
        lines.append('  printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for fname, _ in cls._fields_:
            lines.append('  printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, fname, cname, fname))
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I" + os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    for cls, cname in pairs:
        assert int(out[cname + " size"]) == ctypes.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert int(out["%s.%s" % (cname, fname)]) == getattr(cls, fname).offset, (cname, fname)
