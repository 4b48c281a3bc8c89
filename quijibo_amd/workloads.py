"""Synthetic workloads: seeded root sets and grids for the benchmark configurations.

Roots are quaternions in the reference's slot order (w, x, y, z). A voxel at
spatial (X, Y, Z) iterates from QUATERNION(X, Y, Z, slice)
(OPTIMIZE_3D.cpp:1317), so spatial coordinates fill slots w, x, y and the
"4th" coordinate of a root is slot z. Roots of the drivers are built the same
way, QUATERNION(center[0], center[1], center[2], 0) (cubature.cpp:410,421).

No reference input data exists (the paper's .o3d files are missing blobs), so
every workload is synthetic, generated with numpy.random.default_rng(seed).
"""
from dataclasses import dataclass, field

import numpy as np


@dataclass
class RootSet:
    roots: np.ndarray           # n x 4 float64, (w, x, y, z)
    powers: np.ndarray          # n float64
    power_scalar: float = 1.0

    @property
    def n(self):
        return int(self.roots.shape[0])


@dataclass
class Workload:
    name: str
    res: tuple
    top: RootSet
    bottom: RootSet
    max_iter: int
    escape: float = 2000.0
    exp_scaling: float = 1.0
    slice: float = 0.0
    center: tuple = (0.0, 0.0, 0.0)
    lengths: tuple = (3.0, 3.0, 3.0)
    band: float = 0.05
    extra: dict = field(default_factory=dict)

    @property
    def voxels(self):
        return int(self.res[0]) * int(self.res[1]) * int(self.res[2])


def _empty():
    return RootSet(np.zeros((0, 4)), np.zeros(0))


def _shell_roots(rng, n, r0, r1):
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # volume-uniform radius in [r0, r1]
    u = rng.uniform(size=n)
    r = np.cbrt(r0 ** 3 + u * (r1 ** 3 - r0 ** 3))
    xyz = d * r[:, None]
    return np.concatenate([xyz, np.zeros((n, 1))], axis=1)


def config1(res=64):
    """8 top roots in a shell of radius 0.5-1.0, p=1, maxIter 3, escape 1e3, no bottom."""
    rng = np.random.default_rng(1001)
    top = RootSet(_shell_roots(rng, 8, 0.5, 1.0), np.ones(8))
    return Workload("c1_plain_fp64_64", (res, res, res), top, _empty(), max_iter=3, escape=1e3)


def _gauss_roots(rng, n, sigma, wspread):
    xyz = rng.normal(scale=sigma, size=(n, 3))
    w = rng.uniform(-wspread, wspread, size=(n, 1))
    return np.concatenate([xyz, w], axis=1)


def config2(res=256):
    """32 top + 32 bottom Gaussian roots (sigma 0.6, 4th ~ U[-0.1,0.1]), p ~ U{1,2}, maxIter 4."""
    rng = np.random.default_rng(1002)
    top = RootSet(_gauss_roots(rng, 32, 0.6, 0.1), rng.integers(1, 3, size=32).astype(np.float64))
    bottom = RootSet(_gauss_roots(rng, 32, 0.6, 0.1), rng.integers(1, 3, size=32).astype(np.float64))
    return Workload("c2_rational_32x32_%d" % res, (res, res, res), top, bottom, max_iter=4)


def headline(res=512):
    """The 512^3 headline: the config-2 roots and iteration count at 512^3."""
    w = config2(res)
    w.name = "c2_rational_32x32_%d" % res
    return w


def _torus_roots(rng, n, R, r, jitter):
    u = rng.uniform(0, 2 * np.pi, size=n)
    v = rng.uniform(0, 2 * np.pi, size=n)
    xyz = np.stack([(R + r * np.cos(v)) * np.cos(u), (R + r * np.cos(v)) * np.sin(u), r * np.sin(v)], axis=1)
    xyz += rng.normal(scale=jitter, size=xyz.shape)
    return np.concatenate([xyz, np.zeros((n, 1))], axis=1)


def config3(res=1024, tile=256):
    """128 top + 128 bottom roots on a torus (R=1, r=0.35, jitter 0.02), maxIter 3, 1024^3 in 256^3 tiles."""
    rng = np.random.default_rng(1003)
    top = RootSet(_torus_roots(rng, 128, 1.0, 0.35, 0.02), np.ones(128))
    bottom = RootSet(_torus_roots(rng, 128, 1.0, 0.35, 0.02), np.ones(128))
    w = Workload("c3_torus_128x128_%d" % res, (res, res, res), top, bottom, max_iter=3)
    w.extra["tile"] = tile
    return w


def tile_grid(w, tz, ty, tx, tile):
    """Centre and lengths of one tile of a tiled grid, so that its cell centres
    are exactly the corresponding cells of the full grid (same dx)."""
    res = np.array(w.res, dtype=np.float64)
    L = np.array(w.lengths, dtype=np.float64)
    dx = L / res
    lo = np.array(w.center) - 0.5 * L
    idx = np.array([tx, ty, tz], dtype=np.float64)
    tl = dx * tile
    c = lo + idx * tl + 0.5 * tl
    return tuple(c), tuple(tl)


def config4_requests(n=2000, rate=200.0, seed=1004):
    """Serving stream: Poisson arrivals, edge in {64,128,256,512}, roots 8-256 split
    half/half (Gaussian sigma 0.7), maxIter 3-5."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, size=n)
    t = np.cumsum(gaps) - gaps[0]
    reqs = []
    for i in range(n):
        edge = int(rng.choice([64, 128, 256, 512]))
        nr = int(rng.integers(8, 257))
        nt = nr // 2
        nb = nr - nt
        top = RootSet(_gauss_roots(rng, nt, 0.7, 0.0), np.ones(nt))
        bottom = RootSet(_gauss_roots(rng, nb, 0.7, 0.0), np.ones(nb))
        mi = int(rng.integers(3, 6))
        reqs.append((float(t[i]), Workload("req%d" % i, (edge, edge, edge), top, bottom, max_iter=mi)))
    return reqs


def superquadric_points(n, rng, e1=0.5, e2=2.0, axes=(1.0, 0.8, 0.6)):
    """Points on the superquadric |x/a|^(2/e2) + |y/b|^(2/e2))^(e2/e1) + |z/c|^(2/e1) = 1
    via the signed-power parametrisation, plus outward normals."""
    eta = rng.uniform(-np.pi / 2, np.pi / 2, size=n)
    om = rng.uniform(-np.pi, np.pi, size=n)

    def sp(v, e):
        return np.sign(v) * np.abs(v) ** e

    a, b, c = axes
    x = a * sp(np.cos(eta), e1) * sp(np.cos(om), e2)
    y = b * sp(np.cos(eta), e1) * sp(np.sin(om), e2)
    z = c * sp(np.sin(eta), e1)
    pts = np.stack([x, y, z], axis=1)
    # normal from the implicit gradient
    def g(v, s, e):
        return (2.0 / e) * np.sign(v) * np.abs(v / s) ** (2.0 / e - 1.0) / s
    fx = (np.abs(x / a) ** (2 / e2) + np.abs(y / b) ** (2 / e2))
    k = (e2 / e1) * np.where(fx > 0, fx, 1.0) ** (e2 / e1 - 1.0)
    nx = k * g(x, a, e2)
    ny = k * g(y, b, e2)
    nz = g(z, c, e1)
    nrm = np.stack([nx, ny, nz], axis=1)
    nl = np.linalg.norm(nrm, axis=1, keepdims=True)
    nrm = np.where(nl > 0, nrm / np.where(nl > 0, nl, 1.0), pts / np.linalg.norm(pts, axis=1, keepdims=True))
    return pts, nrm


def config5(n_surface=20000, n_exterior=20000, max_iter=3):
    """Differentiation: 64 top + 64 bottom roots; objective
    J = sum F^2 over surface points + 0.1 * mean exp(-F) over exterior points at distance 0.2."""
    rng = np.random.default_rng(1005)
    top = RootSet(_gauss_roots(rng, 64, 0.6, 0.1), np.ones(64))
    bottom = RootSet(_gauss_roots(rng, 64, 0.6, 0.1), np.ones(64))
    surf, _ = superquadric_points(n_surface, rng)
    ext0, nrm = superquadric_points(n_exterior, rng)
    ext = ext0 + 0.2 * nrm
    w = Workload("c5_grad_64x64", (0, 0, 0), top, bottom, max_iter=max_iter)
    w.extra.update(surface=surf, exterior=ext)
    return w
