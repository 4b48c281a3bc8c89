"""Pins the oracle's marching cubes (TRIANGLE_MESH.cpp:989-1114) and both copies of
the case table: against the reference's own table file where it is present, and
by topology on closed surfaces and single-cell cases everywhere."""
import os
import re

import numpy as np
import pytest

import oracle as O

REF_TABLE = "/root/reference/src/geometry/MARCHING_CUBES_TRIANGLES.include"
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# reference edge (1..12) -> (corner offset, axis), TRIANGLE_MESH.cpp:1504-1563
REF_EDGES = {1: ((0, 0, 0), 1), 2: ((0, 0, 0), 2), 3: ((0, 0, 1), 1), 4: ((0, 1, 0), 2), 5: ((1, 0, 0), 1),
             6: ((1, 0, 0), 2), 7: ((1, 0, 1), 1), 8: ((1, 1, 0), 2), 9: ((0, 1, 0), 0), 10: ((0, 0, 0), 0),
             11: ((0, 0, 1), 0), 12: ((0, 1, 1), 0)}


def _load_c_table(path, name):
    txt = open(path).read()
    body = txt[txt.index("{", txt.index(name)):]
    rows = re.findall(r"\{([-0-9,\s]+)\}", body)[:256]
    return [[int(v) for v in r.split(",")] for r in rows]


def oracle_table():
    rows = _load_c_table(os.path.join(ROOT, "oracle", "qo_mc_table.h"), "QO_MC_TRIS")
    return [[tuple(r[i:i + 3]) for i in range(0, 15, 3) if r[i]] for r in rows]


def product_table():
    rows = _load_c_table(os.path.join(ROOT, "quijibo_amd", "csrc", "qj_mc_table.h"), "QJ_MC_TRI_INIT")
    return [[tuple(r[i:i + 3]) for i in range(0, 15, 3) if r[i] >= 0] for r in rows]


def mine_edge(e):
    """product edge id -> (corner offset, axis)"""
    if e < 4:
        return (0, e & 1, e >> 1), 0
    if e < 8:
        return (e & 1, 0, (e >> 1) & 1), 1
    return (e & 1, (e >> 1) & 1, 0), 2


def test_tables_agree_with_each_other():
    ot, pt = oracle_table(), product_table()
    assert sum(len(r) for r in ot) == 819
    for c in range(256):
        a = [tuple(REF_EDGES[e] for e in t) for t in ot[c]]
        b = [tuple(mine_edge(e) for e in t) for t in pt[c]]
        assert a == b, c


@pytest.mark.skipif(not os.path.exists(REF_TABLE), reason="reference sources not mounted")
def test_oracle_table_is_the_reference_table():
    src = open(REF_TABLE).read()
    parts = re.split(r"case (\d+):", src)
    ref = {}
    for i in range(1, len(parts), 2):
        ref[int(parts[i])] = [tuple(int(v) for v in t) for t in
                              re.findall(r"addVertexPairTriangle\((\d+),(\d+),(\d+),index\)", parts[i + 1])]
    assert sorted(ref) == list(range(256))
    ot = oracle_table()
    for c in range(256):
        assert ot[c] == ref[c], c


def test_oracle_table_matches_the_reference_digest():
    """The same pin without the reference sources: the table's serialisation
    against the SHA-256 that tests/golden/make_mc_table_digest.py took of the
    reference's table file."""
    import hashlib
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from make_mc_table_digest import serialise
    want = open(os.path.join(HERE, "golden", "mc_table_digest.txt")).read().split()[0]
    got = hashlib.sha256(serialise(oracle_table()).encode()).hexdigest()
    assert got == want
    # one changed triangle changes the digest
    t = [list(r) for r in oracle_table()]
    t[158] = t[158] + [(10, 5, 6)]
    assert hashlib.sha256(serialise(t).encode()).hexdigest() != want


def _crossing_edges(flag):
    out = set()
    for e, (o, a) in REF_EDGES.items():
        o2 = list(o); o2[a] += 1
        b0 = (flag >> (4 * o[0] + 2 * o[1] + o[2])) & 1
        b1 = (flag >> (4 * o2[0] + 2 * o2[1] + o2[2])) & 1
        if b0 != b1:
            out.add(e)
    return out


def test_table_rows_use_exactly_the_crossing_edges_except_158():
    ot = oracle_table()
    for c in range(256):
        used = {e for t in ot[c] for e in t}
        if c == 158:
            assert _crossing_edges(c) - used == {5, 6, 10}
        else:
            assert used == _crossing_edges(c), c
    assert {5, 6, 10} <= {e for t in ot[97] for e in t}


def _corner_field(flag):
    """2x2x2 field whose single cell has case `flag`; values +-1 with a tilt."""
    f = np.empty(8)
    for z in range(2):
        for y in range(2):
            for x in range(2):
                bit = 4 * x + 2 * y + z
                f[x + 2 * y + 4 * z] = (1.0 if (flag >> bit) & 1 else -1.0) * (1 + 0.1 * bit)
    return f


@pytest.mark.parametrize("flag", range(256))
def test_single_cell_cases(flag):
    g = O.make_grid((2, 2, 2), (0, 0, 0), (2, 2, 2))
    v, t = O.marching_cubes(g, _corner_field(flag))
    used = {e for tri in oracle_table()[flag] for e in tri}
    assert len(v) == len(used)
    assert len(t) == len(oracle_table()[flag])
    # every vertex lies on its edge between two cell centres (coordinates in [-0.5, 0.5])
    assert np.all(np.abs(v) <= 0.5 + 1e-15)


def _sphere_field(n, r=0.71, c=(0.03, -0.02, 0.01)):
    g = O.make_grid((n, n, n), (0, 0, 0), (2, 2, 2))
    dx = 2.0 / n
    ax = -1 + (np.arange(n) + 0.5) * dx
    Z, Y, X = np.meshgrid(ax, ax, ax, indexing="ij")
    f = r - np.sqrt((X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2)
    return g, f.reshape(-1)


@pytest.mark.parametrize("n", [12, 33])
def test_sphere_is_a_closed_genus0_surface(n):
    g, f = _sphere_field(n)
    v, t = O.marching_cubes(g, f)
    assert len(t) > 0
    # vertex count = sign-changing grid edges (float test), case 158 cannot occur on a sphere
    F = f.reshape(n, n, n) > 0
    n_edges = (F[:, :, 1:] != F[:, :, :-1]).sum() + (F[:, 1:, :] != F[:, :-1, :]).sum() + \
              (F[1:, :, :] != F[:-1, :, :]).sum()
    assert len(v) == n_edges
    edges = {}
    for tri in t:
        for a, b in ((tri[0], tri[1]), (tri[1], tri[2]), (tri[2], tri[0])):
            k = (min(a, b), max(a, b))
            edges[k] = edges.get(k, 0) + 1
    assert all(c == 2 for c in edges.values())  # closed 2-manifold
    assert len(v) - len(edges) + len(t) == 2  # Euler characteristic of a sphere
    rad = np.linalg.norm(v - np.array([0.03, -0.02, 0.01]), axis=1)
    assert np.all(np.abs(rad - 0.71) < 2.0 / n)


def test_vertex_order_is_edge_map_order():
    """Vertices come in std::map<pair<int,int>> order: (lower, upper) linear index."""
    n = 9
    g, f = _sphere_field(n)
    v, t = O.marching_cubes(g, f)
    F = f.reshape(n, n, n) > 0
    keys = []
    for z in range(n):
        for y in range(n):
            for x in range(n):
                p = x + n * y + n * n * z
                for ax, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    if x + dx < n and y + dy < n and z + dz < n and F[z, y, x] != F[z + dz, y + dy, x + dx]:
                        keys.append((p, p + dx + n * dy + n * n * dz))
    assert keys == sorted(keys) and len(keys) == len(v)
    for k, (a, b) in enumerate(keys):
        za, ya, xa = a // (n * n), (a // n) % n, a % n
        zb, yb, xb = b // (n * n), (b // n) % n, b % n
        pa = O.cell_center(g, xa, ya, za)
        pb = O.cell_center(g, xb, yb, zb)
        d0, d1 = f[a], f[b]
        assert np.allclose(v[k], pa + d0 / (d0 - d1) * (pb - pa), rtol=0, atol=1e-15)


def test_float_sign_predicate():
    """(float)v > 0: tiny positives that round to 0 in float and NaN count as outside."""
    g = O.make_grid((2, 2, 2), (0, 0, 0), (2, 2, 2))
    f = -np.ones(8)
    f[0] = 1e-50  # > 0 in double, == 0 in float
    v, t = O.marching_cubes(g, f)
    assert len(v) == 0 and len(t) == 0
    f[0] = np.nan
    v, t = O.marching_cubes(g, f)
    assert len(v) == 0
    f[0] = 1e-30  # positive in float: 3 vertices, but they coincide -> degenerate, dropped
    v, t = O.marching_cubes(g, f)
    assert len(v) == 3 and len(t) == 0
    f[0] = 0.5
    v, t = O.marching_cubes(g, f)
    assert len(v) == 3 and len(t) == 1


def test_nonlinear_mesher_sphere_closed_form():
    """One root, e = 1, one iteration: F = log(c |q - r|), so the surface is the
    sphere |q - r| = 1/c. Bisection vertices lie within edge/2^7 of it; the
    topology equals the linear mesher's on the same field."""
    n = 24
    g = O.make_grid((n, n, n), (0, 0, 0), (3, 3, 3))
    r = np.array([[0.11, -0.07, 0.05, 0.0]])
    c = 1.0 / 0.9
    top = O.poly(r)
    empty = O.poly(np.zeros((0, 4)))
    p = O.make_params(1, 2000.0, c, 0.0, O.BLACK_HOLE)
    v, t, bad = O.nonlinear_mc(g, top, empty, p)
    assert bad == 0 and len(t) > 500
    rad = np.linalg.norm(v - r[0, :3], axis=1)
    dx = 3.0 / n
    assert np.all(np.abs(rad - 0.9) <= dx / 2 ** 7 * 1.0001)
    f = O.field_map(g, top, empty, p)
    lv, lt = O.marching_cubes(g, f)
    assert np.array_equal(lt, t) and len(lv) == len(v)
    assert np.max(np.abs(np.linalg.norm(lv - r[0, :3], axis=1) - 0.9)) > dx / 2 ** 7  # linear is worse


def zero_iterate_case():
    """A top root (power 3) 1e-150 from the midpoint of one x-edge, in the slice
    component: at that midpoint the first iterate underflows to exactly 0. The
    Lagrange mesher's value is then FLT_MAX (TRIANGLE_MESH.cpp:1733-1742) and
    bisection keeps the negative half; the factored mesher's value is -inf
    (:1793-1794) and it keeps the positive half."""
    g = O.make_grid((6, 6, 6), (0.0, 0.0, 0.0), (0.6, 0.6, 0.6))
    p0 = O.cell_center(g, 2, 2, 2)
    p1 = O.cell_center(g, 3, 2, 2)
    m = (p0 + p1) * 0.5
    top = np.array([[m[0], m[1], m[2], -1e-150]])
    bot = np.array([[p1[0] + 0.01, p1[1], p1[2], 0.0]])
    return g, p0, p1, m, top, bot, float(np.exp(5.5))


def edge_vertex(v, p0, p1):
    sel = v[(v[:, 1] == p0[1]) & (v[:, 2] == p0[2]) & (v[:, 0] > p0[0]) & (v[:, 0] < p1[0])]
    assert len(sel) == 1
    return sel[0]


def test_nonlinear_mesher_iterate_on_zero():
    g, p0, p1, m, top, bot, c = zero_iterate_case()
    F = O.field_map(g, O.poly(top, [3.0]), O.poly(bot, [1.0]), O.make_params(1, 2000.0, c, 0.0)).reshape(6, 6, 6)
    assert F[2, 2, 2] < 0 < F[2, 2, 3]  # the edge straddles the surface
    lag = O.nonlinear_mc(g, O.poly(top, [3.0]), O.poly(bot, [1.0]),
                         O.make_params(1, 2000.0, c, 0.0, O.BLACK_HOLE | O.LAGRANGE))
    fac = O.nonlinear_mc(g, O.poly(top, [3.0]), O.poly(bot, [1.0]), O.make_params(1, 2000.0, c, 0.0, O.BLACK_HOLE))
    assert lag[2] == 0 and fac[2] == 0
    assert p0[0] < edge_vertex(lag[0], p0, p1)[0] <= m[0]
    assert m[0] <= edge_vertex(fac[0], p0, p1)[0] < p1[0]
