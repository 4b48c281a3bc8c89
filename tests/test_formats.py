"""File formats: the Python writers produce the reference's byte layouts
(restated here with struct from the reference's write functions), the C++ host
layer reads what Python writes and writes the same bytes back, and both read
what the other wrote."""
import gzip
import os
import struct
import subprocess

import numpy as np
import pytest

from quijibo_amd import formats as FM

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "cpp", "_build", "host_layer_driver")


def _sample():
    top = FM.Polynomial4D([[0.5, -0.25, 0.125, 0.0], [-0.75, 0.5, 0.3, 0.01]], [1.0, 2.0], 1.25)
    bottom = FM.Polynomial4D([[0.1, 0.2, 0.3, 0.4]])
    f = FM.Field3D((5, 4, 3), (0.1, -0.2, 0.3), (3.0, 2.0, 1.5), np.arange(60) * 0.5 - 7.0)
    d = FM.Field3D((5, 4, 3), (0.1, -0.2, 0.3), (3.0, 2.0, 1.5), np.arange(60) ** 2 * 1e-3)
    return FM.O3D((5, 4, 3), (0.1, -0.2, 0.3), (3.0, 2.0, 1.5), top, bottom, 0.75, 0, f, d, None)


def test_field_layout(tmp_path):
    f = _sample().fractal
    p = tmp_path / "f.bin"
    FM.write_field(p, f)
    # FIELD_3D::write (FIELD_3D.cpp:4371-4395): 3 ints, center, lengths, cells as doubles
    want = struct.pack("<iii", 5, 4, 3) + struct.pack("<ddd", 0.1, -0.2, 0.3) + struct.pack("<ddd", 3.0, 2.0, 1.5)
    want += b"".join(struct.pack("<d", v) for v in f.data)
    assert p.read_bytes() == want
    g = FM.read_field(p)
    assert g.res == (5, 4, 3) and np.array_equal(g.data, f.data)
    FM.write_field(tmp_path / "f.bin.gz", f)
    assert gzip.decompress((tmp_path / "f.bin.gz").read_bytes()) == want
    assert np.array_equal(FM.read_field(tmp_path / "f.bin.gz").data, f.data)


def test_polynomial_layout(tmp_path):
    p = _sample().top
    FM.write_polynomial(tmp_path / "p.bin", p)
    # POLYNOMIAL_4D::write (POLYNOMIAL_4D.cpp:650-685)
    want = struct.pack("<i", 2) + struct.pack("<iii", 0, 0, 0) + struct.pack("<i", 2)
    want += struct.pack("<8d", 0.5, -0.25, 0.125, 0.0, -0.75, 0.5, 0.3, 0.01)
    want += struct.pack("<i", 2) + struct.pack("<2d", 1.0, 2.0) + struct.pack("<d", 1.25)
    assert (tmp_path / "p.bin").read_bytes() == want
    # a file with expanded coefficient arrays (as the reference writes them) reads the same
    coeffs = struct.pack("<i", 3) + struct.pack("<12d", *range(12))
    full = struct.pack("<i", 2) + coeffs + struct.pack("<i", 1) + struct.pack("<4d", 9, 9, 9, 9) + struct.pack("<i", 0)
    full += want[4 + 12:]
    (tmp_path / "q.bin").write_bytes(full)
    q = FM.read_polynomial(tmp_path / "q.bin")
    assert np.array_equal(q.roots, p.roots) and np.array_equal(q.powers, p.powers) and q.power_scalar == 1.25


def test_o3d_layout(tmp_path):
    s = _sample()
    FM.write_o3d(tmp_path / "s.o3d", s)
    raw = (tmp_path / "s.o3d").read_bytes()
    # OPTIMIZE_3D::write (OPTIMIZE_3D.cpp:3254-3294)
    head = struct.pack("<iii", 5, 4, 3) + struct.pack("<6d", 0.1, -0.2, 0.3, 3.0, 2.0, 1.5)
    assert raw.startswith(head)
    FM.write_polynomial(tmp_path / "t.bin", s.top)
    FM.write_polynomial(tmp_path / "b.bin", s.bottom)
    pt, pb = (tmp_path / "t.bin").read_bytes(), (tmp_path / "b.bin").read_bytes()
    o = len(head)
    assert raw[o:o + len(pt)] == pt and raw[o + len(pt):o + len(pt) + len(pb)] == pb
    o += len(pt) + len(pb)
    assert struct.unpack("<di", raw[o:o + 12]) == (0.75, 0)
    r = FM.read_o3d(tmp_path / "s.o3d")
    assert np.array_equal(r.fractal.data, s.fractal.data) and np.array_equal(r.distance.data, s.distance.data)
    assert r.curvature.data.size == 0 and r.exp_scaling == 0.75
    with pytest.raises(ValueError):
        (tmp_path / "short.o3d").write_bytes(raw[:100])
        FM.read_o3d(tmp_path / "short.o3d")


def test_obj_round_trip(tmp_path):
    v = np.array([[0.1, 0.2, 0.3], [1.0 / 3, -2.0, 5e-9], [0.0, 1.0, 0.0]])
    t = np.array([[0, 1, 2]])
    FM.write_obj(tmp_path / "m.obj", v, t)
    lines = (tmp_path / "m.obj").read_text().splitlines()
    assert lines[0] == "v 0.1000000000000000 0.2000000000000000 0.3000000000000000"
    assert lines[-1] == "f 1 2 3"
    rv, rt = FM.read_obj(tmp_path / "m.obj")
    assert np.allclose(rv, v, atol=1e-16) and np.array_equal(rt, t)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="host_layer_driver not built")
def test_cpp_layer_io(tmp_path):
    r = subprocess.run([DRIVER, "io", str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    # what C++ wrote, Python reads
    f = FM.read_field(tmp_path / "field.bin")
    assert f.res == (5, 4, 3) and f.center == (0.1, -0.2, 0.3)
    i = np.arange(60, dtype=np.float64)
    assert np.array_equal(f.data, 0.5 * i - 7.0 + 1e-3 * (i * i))
    assert np.array_equal(FM.read_field(tmp_path / "field.bin.gz").data, f.data)
    s = FM.read_o3d(tmp_path / "state.o3d")
    assert s.exp_scaling == 0.75 and s.top.power_scalar == 1.25 and s.bottom.roots.tolist() == [[0.1, 0.2, 0.3, 0.4]]
    assert s.distance.data[0] == 42.0 and s.curvature.data.size == 0


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="host_layer_driver not built")
def test_cpp_layer_reads_and_rewrites_python_o3d(tmp_path):
    FM.write_o3d(tmp_path / "a.o3d", _sample())
    r = subprocess.run([DRIVER, "copy", str(tmp_path / "a.o3d"), str(tmp_path / "b.o3d")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert (tmp_path / "a.o3d").read_bytes() == (tmp_path / "b.o3d").read_bytes()
    (tmp_path / "c.o3d").write_bytes((tmp_path / "a.o3d").read_bytes()[:200])
    r = subprocess.run([DRIVER, "copy", str(tmp_path / "c.o3d"), str(tmp_path / "d.o3d")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 2 and "too short" in r.stderr


def test_edge_cache_layout(tmp_path):
    edges = np.array([[1, 2, 3, 0], [1, 2, 3, 2], [4, 0, 3, 1]], dtype=np.int32)
    cells = np.array([[1, 2, 3, 17], [4, 0, 2, 254]], dtype=np.int32)
    FM.write_edge_cache(tmp_path / "c.cache", edges, cells)
    # writeEdgeCacheHuge (TRIANGLE_MESH.cpp:531-583): count, (first xyz, second xyz)...,
    # count, (flag, xyz)...
    want = struct.pack("<i", 3) + struct.pack("<6i", 1, 2, 3, 2, 2, 3) + struct.pack("<6i", 1, 2, 3, 1, 2, 4)
    want += struct.pack("<6i", 4, 0, 3, 4, 1, 3)
    want += struct.pack("<i", 2) + struct.pack("<4i", 17, 1, 2, 3) + struct.pack("<4i", 254, 4, 0, 2)
    assert (tmp_path / "c.cache").read_bytes() == want
    e, c = FM.read_edge_cache(tmp_path / "c.cache")
    assert np.array_equal(e, edges) and np.array_equal(c, cells)
    bad = struct.pack("<i", 1) + struct.pack("<6i", 0, 0, 0, 1, 1, 0) + struct.pack("<i", 0)
    (tmp_path / "b.cache").write_bytes(bad)
    with pytest.raises(ValueError):
        FM.read_edge_cache(tmp_path / "b.cache")
