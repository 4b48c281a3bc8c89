"""Build libquijibo.so (HIP kernels for gfx950 + the C ABI) in-tree with hipcc.

Usage: python -m quijibo_amd.build   (or __graft_entry__.build())
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libquijibo.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("QJ_OFFLOAD_ARCH", "gfx950")

SOURCES = ["csrc/qj_api.cpp", "csrc/qj_field.hip", "csrc/qj_mc.hip", "csrc/qj_scan.hip", "csrc/qj_grad.hip", "csrc/qj_fieldops.hip",
           "csrc/qj_server.cpp", "cpp/quijibo.cpp", "cpp/toupee.cpp", "cpp/tooth_optimizer.cpp"]

# Packed FP32 (v_pk_*_f32) is disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-Wall", "-Wno-unused-function",
          "-Wno-unused-variable", "-Wno-unused-but-set-variable", "-Wno-unused-value", "-Wno-unused-result", "-I" + os.path.join(HERE, "..", "include")]


# The FP32 field kernels branch per root on the exponent kind with the
# quaternion product inside each branch; SimplifyCFG's common-code sinking
# would merge those products behind register copies.
FILE_FLAGS = {"csrc/qj_field.hip": ["-mllvm", "-simplifycfg-sink-common=false", "-mllvm", "-structurizecfg-skip-uniform-regions"]}


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = ([HIPCC, "-x", "hip"] + COMMON + DEVICE_FLAGS + FILE_FLAGS.get(src, []) +
           ["-c", os.path.join(HERE, src), "-o", obj])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, " ".join(cmd), r.stderr[-6000:]))
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, err


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for obj, err in results:
        if verbose and err:
            print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + [o for o, _ in results] + ["-lpthread", "-lz"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    build_driver()
    return LIB


DRIVER_SRC = os.path.join(HERE, "..", "tests", "cpp", "host_layer_driver.cpp")
DRIVER = os.path.join(HERE, "..", "tests", "cpp", "_build", "host_layer_driver")


def build_driver():
    """The C++ host-layer test program (a plain host executable linked to libquijibo.so)."""
    os.makedirs(os.path.dirname(DRIVER), exist_ok=True)
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-I" + os.path.join(HERE, "..", "include"),
           "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__", DRIVER_SRC, "-o", DRIVER + ".tmp", "-L" + HERE,
           "-lquijibo", "-Wl,-rpath," + os.path.relpath(HERE, os.path.dirname(DRIVER)).join(["$ORIGIN/", ""]),
           "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("driver build failed:\n%s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    os.replace(DRIVER + ".tmp", DRIVER)
    return DRIVER


if __name__ == "__main__":
    print(build(verbose=True))
