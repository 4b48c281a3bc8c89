"""Binary and text formats of the reference, read and written with numpy.

Layouts (little-endian, int = int32, every real stored as a double):
  FIELD_3D      (FIELD_3D.cpp:4371-4450): int xRes, yRes, zRes; center[3];
                lengths[3]; xRes*yRes*zRes cells, x fastest. A file name
                ending in "gz" is gzip-compressed (FIELD_3D.cpp:280-390).
  POLYNOMIAL_4D (POLYNOMIAL_4D.cpp:650-728): int totalRoots; three arrays of
                quaternions (int count + count*4 doubles: expanded
                coefficients, derivatives, second derivatives); int count +
                roots (w,x,y,z); int count + powers; double powerScalar.
                Reading keeps only roots, powers and powerScalar; writing
                stores the three expanded arrays empty.
  .o3d          (OPTIMIZE_3D.cpp:3254-3346): int xRes, yRes, zRes; center[3];
                lengths[3]; top and bottom POLYNOMIAL_4D; double expScaling;
                int history size (its entries are never written); the fractal,
                distance and curvature FIELD_3Ds.
  OBJ           (TRIANGLE_MESH.cpp:2762-2800): "v %.16f %.16f %.16f" lines,
                then "f i j k" with 1-based indices.
"""
import gzip
import struct
from dataclasses import dataclass

import numpy as np


@dataclass
class Field3D:
    res: tuple                       # (xRes, yRes, zRes)
    center: tuple = (0.0, 0.0, 0.0)
    lengths: tuple = (1.0, 1.0, 1.0)
    data: np.ndarray = None          # float64, xRes*yRes*zRes, x fastest

    def __post_init__(self):
        n = int(np.prod(self.res)) if min(self.res) >= 0 else 0
        if self.data is None:
            self.data = np.zeros(n)
        self.data = np.ascontiguousarray(self.data, dtype="<f8").reshape(-1)
        if self.data.size != n:
            raise ValueError("field data has %d cells, resolution needs %d" % (self.data.size, n))


@dataclass
class Polynomial4D:
    roots: np.ndarray                # n x 4 (w, x, y, z)
    powers: np.ndarray = None
    power_scalar: float = 1.0

    def __post_init__(self):
        self.roots = np.asarray(self.roots, dtype=np.float64).reshape(-1, 4)
        if self.powers is None:
            self.powers = np.ones(self.roots.shape[0])
        self.powers = np.asarray(self.powers, dtype=np.float64).reshape(-1)
        if self.powers.size != self.roots.shape[0]:
            raise ValueError("one power per root")


@dataclass
class O3D:
    res: tuple
    center: tuple
    lengths: tuple
    top: Polynomial4D
    bottom: Polynomial4D
    exp_scaling: float = 1.0
    history_size: int = 0
    fractal: Field3D = None
    distance: Field3D = None
    curvature: Field3D = None

    def __post_init__(self):
        empty = Field3D((0, 0, 0))
        self.fractal = self.fractal or empty
        self.distance = self.distance or empty
        self.curvature = self.curvature or empty


class _In:
    def __init__(self, buf):
        self.b = memoryview(buf)
        self.o = 0

    def take(self, n):
        if self.o + n > len(self.b):
            raise ValueError("file too short: need %d bytes at offset %d of %d" % (n, self.o, len(self.b)))
        v = self.b[self.o:self.o + n]
        self.o += n
        return v

    def i32(self):
        return struct.unpack("<i", self.take(4))[0]

    def f64(self, n=None):
        if n is None:
            return struct.unpack("<d", self.take(8))[0]
        return np.frombuffer(self.take(8 * n), dtype="<f8").copy()


def _pack_field(f):
    out = struct.pack("<3i", *[int(v) for v in f.res]) + struct.pack("<3d", *f.center) + struct.pack("<3d", *f.lengths)
    if f.data.size:
        out += np.ascontiguousarray(f.data, dtype="<f8").tobytes()
    return out


def _unpack_field(r):
    res = (r.i32(), r.i32(), r.i32())
    c = tuple(r.f64(3))
    l = tuple(r.f64(3))
    if min(res) < 0:
        return Field3D((0, 0, 0), c, l)
    n = res[0] * res[1] * res[2]
    return Field3D(res, c, l, r.f64(n) if n else np.zeros(0))


def _pack_poly(p):
    n = p.roots.shape[0]
    out = struct.pack("<i", n) + struct.pack("<3i", 0, 0, 0) + struct.pack("<i", n)
    out += np.ascontiguousarray(p.roots, dtype="<f8").tobytes()
    out += struct.pack("<i", n) + np.ascontiguousarray(p.powers, dtype="<f8").tobytes()
    return out + struct.pack("<d", p.power_scalar)


def _unpack_poly(r):
    total = r.i32()
    if total < 0:
        raise ValueError("negative root count")
    for _ in range(3):
        k = r.i32()
        if k < 0:
            raise ValueError("negative array size")
        r.take(32 * k)
    nr = r.i32()
    if nr != total:
        raise ValueError("root count %d != totalRoots %d" % (nr, total))
    roots = r.f64(4 * nr).reshape(-1, 4)
    r.i32()  # stored power count; the reference reads totalRoots powers
    powers = r.f64(total)
    return Polynomial4D(roots, powers, r.f64())


def _read_bytes(path):
    with open(path, "rb") as fh:
        raw = fh.read()
    return gzip.decompress(raw) if str(path).endswith("gz") else raw


def _write_bytes(path, data):
    if str(path).endswith("gz"):
        data = gzip.compress(data, compresslevel=1)
    with open(path, "wb") as fh:
        fh.write(data)


def write_field(path, f):
    _write_bytes(path, _pack_field(f))


def read_field(path):
    return _unpack_field(_In(_read_bytes(path)))


def write_polynomial(path, p):
    _write_bytes(path, _pack_poly(p))


def read_polynomial(path):
    return _unpack_poly(_In(_read_bytes(path)))


def write_o3d(path, s):
    out = struct.pack("<3i", *[int(v) for v in s.res]) + struct.pack("<3d", *s.center) + struct.pack("<3d", *s.lengths)
    out += _pack_poly(s.top) + _pack_poly(s.bottom) + struct.pack("<d", s.exp_scaling)
    out += struct.pack("<i", int(s.history_size))
    out += _pack_field(s.fractal) + _pack_field(s.distance) + _pack_field(s.curvature)
    with open(path, "wb") as fh:
        fh.write(out)


def read_o3d(path):
    with open(path, "rb") as fh:
        r = _In(fh.read())
    res = (r.i32(), r.i32(), r.i32())
    c = tuple(r.f64(3))
    l = tuple(r.f64(3))
    top = _unpack_poly(r)
    bottom = _unpack_poly(r)
    e = r.f64()
    h = r.i32()
    return O3D(res, c, l, top, bottom, e, h, _unpack_field(r), _unpack_field(r), _unpack_field(r))


def write_obj(path, vertices, triangles):
    with open(path, "w") as fh:
        for v in np.asarray(vertices, dtype=np.float64).reshape(-1, 3):
            fh.write("v %.16f %.16f %.16f\n" % (v[0], v[1], v[2]))
        for t in np.asarray(triangles).reshape(-1, 3):
            fh.write("f %i %i %i\n" % (t[0] + 1, t[1] + 1, t[2] + 1))
    return True


def read_obj(path):
    """Vertices (n x 3 float64) and 0-based triangles (m x 3 int32) of an OBJ file."""
    v, t = [], []
    with open(path) as fh:
        for line in fh:
            if line.startswith("v "):
                v.append([float(x) for x in line.split()[1:4]])
            elif line.startswith("f "):
                t.append([int(x.split("/")[0]) - 1 for x in line.split()[1:4]])
    return np.array(v, dtype=np.float64).reshape(-1, 3), np.array(t, dtype=np.int32).reshape(-1, 3)


def edges_to_triplets(edges):
    """(x, y, z, axis) rows -> the cache's (first xyz, second xyz) grid-point pairs."""
    e = np.asarray(edges, dtype=np.int32).reshape(-1, 4)
    first = e[:, :3]
    second = first + np.eye(3, dtype=np.int32)[e[:, 3]] if len(e) else first.copy()
    return first, second


def write_edge_cache(path, edges, cells):
    """TRIANGLE_MESH::writeEdgeCacheHuge (TRIANGLE_MESH.cpp:531-583): int count, then per
    vertex the two grid points of its edge as 3 + 3 ints (map order); int count, then per
    active cell int flag and the cell's 3 ints."""
    first, second = edges_to_triplets(edges)
    c = np.asarray(cells, dtype=np.int32).reshape(-1, 4)
    out = struct.pack("<i", len(first)) + np.ascontiguousarray(np.c_[first, second], dtype="<i4").tobytes()
    out += struct.pack("<i", len(c)) + np.ascontiguousarray(c[:, [3, 0, 1, 2]], dtype="<i4").tobytes()
    with open(path, "wb") as fh:
        fh.write(out)


def read_edge_cache(path):
    """readEdgeCacheHuge (TRIANGLE_MESH.cpp:431-481) -> (edges n x 4 [x, y, z, axis], cells m x 4 [x, y, z, flag])."""
    with open(path, "rb") as fh:
        r = _In(fh.read())
    n = r.i32()
    if n < 0:
        raise ValueError("negative triplet count")
    t = np.frombuffer(r.take(24 * n), dtype="<i4").reshape(-1, 6).astype(np.int32)
    d = t[:, 3:] - t[:, :3]
    if len(d) and not np.all((d.sum(axis=1) == 1) & (d.min(axis=1) == 0)):
        raise ValueError("a triplet's points are not grid neighbours along +x, +y or +z")
    edges = np.c_[t[:, :3], np.argmax(d, axis=1) if len(d) else np.zeros(0, np.int32)].astype(np.int32)
    m = r.i32()
    if m < 0:
        raise ValueError("negative flag count")
    f = np.frombuffer(r.take(16 * m), dtype="<i4").reshape(-1, 4).astype(np.int32)
    return edges.reshape(-1, 4), f[:, [1, 2, 3, 0]].reshape(-1, 4)
