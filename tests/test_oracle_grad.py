"""Pins the oracle's one-iteration AP score gradients (oracle/qo.c
qo_ap_score_gradients) without trusting its own formulas.

* Top powers and the top power scalar (computeTopGradientAP,
  OPTIMIZE_3D.cpp:8905-9106; computeTopBulkScalarGradientAP, :9250-9330): the
  reference formula is the exact derivative of the score on any grid, because
  d/de (b^e) = b^e log b and b^e commutes with log b. Central differences of
  the oracle score pin them on a general 3D grid.
* Bottom powers and the bottom power scalar (computeBottomGradientAP,
  :10012-10100; computeBottomBulkScalarGradientAP, :10360-10440): the
  reference formula T (-L (P log(b)^-1 / s) R_suf)^-1 does not look like a
  derivative, but at one iteration it is one. The field is
  F = log c + log|T| - log|B| (quaternion norms are multiplicative), and
  <a X b, a Y b> = |a|^2 |b|^2 <X, Y>, so each cell's bottom term reduces to
  dSdF * (-s_b log|q0 - b_j|), the exact derivative. Central differences of
  the oracle score therefore pin them on a general 3D grid.
* On one z-slice at Z = 0 with slice 0 and roots whose j and k parts are 0,
  every quaternion lies in the complex i-plane, the one-iteration field is
      F = log c + s_t sum_i p_i log|q0 - t_i| - s_b sum_j p_j log|q0 - b_j|,
  and every output has a closed form in numpy.
* Off that plane the factors' order matters term by term: single-cell grids
  are checked against the reference's text restated with 2x2 complex
  matrices (explicit ordered products of principal matrix powers, scipy's
  logm). A copy of that restatement with planted mistakes (prefix and suffix
  swapped; 1/s in place of 1/p_j) is shown to fail the difference check.
"""
import numpy as np
import scipy.linalg

import oracle as O

ALPHA = 10.0


def M(q):
    a, b, c, d = q
    return np.array([[a + 1j * b, c + 1j * d], [-c + 1j * d, a - 1j * b]])


def qdot(A, B):
    # the quaternion dot product (QUATERNION.h:137) of two matrix forms
    return float(np.real(np.trace(A.conj().T @ B))) / 2.0


def plane_case(seed=3, nt=5, nb=4):
    """A commuting configuration: grid of one z-slice at Z = 0, roots in the
    (w, x) plane, slice 0."""
    rng = np.random.default_rng(seed)
    top = np.zeros((nt, 4))
    bot = np.zeros((nb, 4))
    top[:, :2] = rng.normal(scale=0.7, size=(nt, 2))
    bot[:, :2] = rng.normal(scale=0.7, size=(nb, 2))
    tp = rng.uniform(0.8, 2.2, nt)
    bp = rng.uniform(0.8, 2.2, nb)
    res = (24, 20, 1)
    center = (0.1, -0.05, 0.0)
    lengths = (3.0, 2.8, 0.5)
    ax = [center[i] - lengths[i] / 2 + (np.arange(res[i]) + 0.5) * lengths[i] / res[i] for i in range(2)]
    Y, X = np.meshgrid(ax[1], ax[0], indexing="ij")
    weight = (1.0 + 0.5 * np.sin(3 * X) * np.cos(2 * Y)).reshape(-1)
    return dict(top=top, bot=bot, tp=tp, bp=bp, st=0.9, sb=1.3, c=0.7, res=res, center=center,
                lengths=lengths, weight=weight)


def plane_closed_form(case, thr=0.0):
    """Score and gradients of the one-iteration AP map from the complex closed
    form; cell centres as in FIELD_3D::cellCenter (FIELD_3D.cpp:818-841)."""
    res, center, lengths = case["res"], case["center"], case["lengths"]
    xs = [(center[i] - lengths[i] * 0.5) + np.arange(res[i]) * (lengths[i] / res[i]) + 0.5 * (lengths[i] / res[i])
          - center[i] + center[i] for i in range(2)]
    Y, X = np.meshgrid(xs[1], xs[0], indexing="ij")
    z = (X + 1j * Y).reshape(-1)
    lt = np.log(np.abs(z[:, None] - (case["top"][:, 0] + 1j * case["top"][:, 1])[None, :]))
    lb = np.log(np.abs(z[:, None] - (case["bot"][:, 0] + 1j * case["bot"][:, 1])[None, :]))
    F = np.log(case["c"]) + case["st"] * lt @ case["tp"] - case["sb"] * lb @ case["bp"]
    w = case["weight"]
    dSdF = -ALPHA * w * (1.0 - np.tanh(ALPHA * (F - thr)) ** 2)
    return dict(field=F, score=float(np.sum(-np.tanh(ALPHA * F) * w)), grad_exp=float(np.sum(dSdF)),
                grad_top=case["st"] * (dSdF @ lt), grad_bottom=-case["sb"] * (dSdF @ lb),
                grad_top_scalar=float(dSdF @ (lt @ case["tp"])),
                grad_bottom_scalar=float(-(dSdF @ (lb @ case["bp"]))))


def oracle_ap(case, top=None, bot=None, tp=None, bp=None, st=None, sb=None, c=None):
    g = O.make_grid(case["res"], case["center"], case["lengths"])
    return O.ap_score_gradients(
        g, O.poly(case["top"] if top is None else top, case["tp"] if tp is None else tp, case["st"] if st is None else st),
        O.poly(case["bot"] if bot is None else bot, case["bp"] if bp is None else bp, case["sb"] if sb is None else sb),
        O.make_params(1, 20.0, case["c"] if c is None else c, 0.0, O.SKIP_ORIGIN | O.AP_REPASS), g,
        case["weight"], ALPHA, 0.0)


def _rel(a, b):
    a = np.atleast_1d(a); b = np.atleast_1d(b)
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def test_plane_case_is_commuting_and_field_matches_closed_form():
    case = plane_case()
    cf = plane_closed_form(case)
    r = oracle_ap(case)
    assert r["overflowed"] == 0
    assert np.allclose(r["field"], cf["field"], rtol=0, atol=1e-12)
    assert abs(r["score"] - cf["score"]) <= 1e-12 * abs(cf["score"]) * len(cf["field"])


def test_oracle_gradients_match_closed_form_on_commuting_plane():
    case = plane_case()
    cf = plane_closed_form(case)
    r = oracle_ap(case)
    for k in ("grad_exp", "grad_top", "grad_bottom", "grad_top_scalar", "grad_bottom_scalar"):
        e = _rel(r[k], cf[k])
        assert e < 1e-10, (k, e, r[k], cf[k])


def test_oracle_bottom_gradients_are_score_derivatives_on_commuting_plane():
    case = plane_case()
    r = oracle_ap(case)
    h = 1e-6
    fd = []
    for j in range(case["bot"].shape[0]):
        s = []
        for sg in (1, -1):
            bp = case["bp"].copy()
            bp[j] += sg * h
            s.append(oracle_ap(case, bp=bp)["score"])
        fd.append((s[0] - s[1]) / (2 * h))
    assert _rel(r["grad_bottom"], np.array(fd)) < 1e-6
    s = [oracle_ap(case, sb=case["sb"] + sg * h)["score"] for sg in (1, -1)]
    fd_sb = (s[0] - s[1]) / (2 * h)
    assert abs(r["grad_bottom_scalar"] - fd_sb) < 1e-6 * abs(fd_sb)
    # chain rule: d/ds_b = sum_j (p_j / s_b) d/dp_j
    assert abs(float(np.dot(fd, case["bp"])) / case["sb"] - fd_sb) < 1e-6 * abs(fd_sb)
    # a planted mistake of the kind these checks must catch: 1/s in place of
    # 1/p_j in the bulk term turns each bulk term into the per-root term
    planted = float(np.sum(fd))
    assert abs(planted - fd_sb) > 1e-2 * abs(fd_sb)


def test_oracle_top_scalar_gradient_is_the_score_derivative_on_a_3d_grid():
    rng = np.random.default_rng(11)
    case = dict(top=rng.normal(scale=0.6, size=(6, 4)), bot=rng.normal(scale=0.6, size=(5, 4)),
                tp=rng.uniform(0.8, 2.0, 6), bp=rng.uniform(0.8, 2.0, 5), st=1.1, sb=0.85, c=0.8,
                res=(11, 10, 9), center=(0.05, 0.0, -0.1), lengths=(2.6, 2.6, 2.6))
    case["top"][:, 3] *= 0.2
    case["bot"][:, 3] *= 0.2
    case["weight"] = rng.uniform(0.5, 2.0, 11 * 10 * 9)
    r = oracle_ap(case)
    assert r["overflowed"] == 0
    h = 1e-6
    s = [oracle_ap(case, st=case["st"] + sg * h)["score"] for sg in (1, -1)]
    fd = (s[0] - s[1]) / (2 * h)
    assert abs(r["grad_top_scalar"] - fd) < 1e-6 * abs(fd), (r["grad_top_scalar"], fd)
    fdt = []
    for i in range(6):
        s = []
        for sg in (1, -1):
            tp = case["tp"].copy()
            tp[i] += sg * h
            s.append(oracle_ap(case, tp=tp)["score"])
        fdt.append((s[0] - s[1]) / (2 * h))
    assert _rel(r["grad_top"], np.array(fdt)) < 1e-6
    # the exact scalar derivative is the power-weighted sum of the per-root terms / s
    assert abs(r["grad_top_scalar"] - float(np.dot(r["grad_top"], case["tp"])) / case["st"]) < 1e-9 * abs(fd)


def _matrix_terms(q0, roots, powers, s):
    bases = [M(q0 - r) for r in roots]
    P = [scipy.linalg.fractional_matrix_power(b, s * p) for b, p in zip(bases, powers)]
    return bases, P


def _prod(ms):
    out = np.eye(2, dtype=complex)
    for m in ms:
        out = out @ m
    return out


def matrix_form_cell(q0, w, case, thr=0.0):
    """The reference's per-cell gradient terms written out with 2x2 complex
    matrices: compute{Top,Bottom}GradientAP (OPTIMIZE_3D.cpp:8942-9000,
    10044-10087) and the bulk-scalar variants (:9310-9317, :10422-10430)."""
    tb, TP = _matrix_terms(q0, case["top"], case["tp"], case["st"])
    bb, BP = _matrix_terms(q0, case["bot"], case["bp"], case["sb"])
    T = _prod(TP)
    B = _prod(BP)
    R = T @ np.linalg.inv(B)
    F = np.log(np.sqrt(qdot(R, R))) + np.log(case["c"])
    dSdF = -ALPHA * w * (1 - np.tanh(ALPHA * (F - thr)) ** 2)
    out = dict(grad_top=[], grad_bottom=[], grad_top_scalar=0.0, grad_bottom_scalar=0.0)
    for i in range(len(TP)):
        L, Rs = _prod(TP[:i]), _prod(TP[i + 1:])
        for bulk in (0, 1):
            mid = (case["tp"][i] if bulk else case["st"]) * TP[i] @ scipy.linalg.logm(tb[i])
            Rd = (L @ mid @ Rs) @ np.linalg.inv(B)
            v = dSdF * qdot(R, Rd) / qdot(R, R)
            if bulk:
                out["grad_top_scalar"] += v
            else:
                out["grad_top"].append(v)
    for j in range(len(BP)):
        L, Rs = _prod(BP[:j]), _prod(BP[j + 1:])
        for bulk in (0, 1):
            mid = (1.0 / (case["bp"][j] if bulk else case["sb"])) * BP[j] @ np.linalg.inv(scipy.linalg.logm(bb[j]))
            Bd = -1.0 * L @ mid @ Rs
            Rd = T @ np.linalg.inv(Bd)
            v = dSdF * qdot(R, Rd) / qdot(R, R)
            if bulk:
                out["grad_bottom_scalar"] += v
            else:
                out["grad_bottom"].append(v)
    return out


def _scaling_near_surface(q0, case, F=0.03):
    """exp scaling that puts the cell's field at F, so the tanh slope is not 0"""
    T = _prod(_matrix_terms(q0, case["top"], case["tp"], case["st"])[1])
    B = _prod(_matrix_terms(q0, case["bot"], case["bp"], case["sb"])[1])
    R = T @ np.linalg.inv(B)
    return float(np.exp(F) / np.sqrt(qdot(R, R)))


def _single_cell(p):
    return dict(res=(1, 1, 1), center=tuple(p), lengths=(0.1, 0.1, 0.1))


def test_oracle_gradients_match_matrix_form_off_the_plane():
    rng = np.random.default_rng(5)
    base = dict(top=rng.normal(scale=0.6, size=(4, 4)), bot=rng.normal(scale=0.6, size=(3, 4)),
                tp=rng.uniform(0.8, 2.0, 4), bp=rng.uniform(0.8, 2.0, 3), st=1.15, sb=0.8, c=0.9)
    slice_ = 0.0
    for _ in range(6):
        case = dict(base, **_single_cell(rng.uniform(-1.0, 1.0, 3)))
        case["weight"] = np.array([1.3])
        g = O.make_grid(case["res"], case["center"], case["lengths"])
        c = O.cell_center(g, 0, 0, 0)
        q0 = np.array([c[0], c[1], c[2], slice_])
        case["c"] = _scaling_near_surface(q0, case)
        mf = matrix_form_cell(q0, 1.3, case)
        r = oracle_ap(case)
        assert r["overflowed"] == 0
        for k in ("grad_top", "grad_bottom", "grad_top_scalar", "grad_bottom_scalar"):
            e = _rel(r[k], np.asarray(mf[k]))
            # the reference's log takes |v| in single precision (VEC3.h:101-103)
            assert e < 1e-5, (k, e, r[k], mf[k])


def test_oracle_bottom_gradients_are_score_derivatives_on_a_3d_grid():
    """Off the plane as well: at one iteration F = log c + log|T| - log|B|, and
    <a X b, a Y b> = |a|^2 |b|^2 <X, Y> for quaternions, so the reference's
    bottom term reduces to dSdF * (-s_b log|q0 - b_j|), the exact derivative."""
    rng = np.random.default_rng(12)
    case = dict(top=rng.normal(scale=0.6, size=(5, 4)), bot=rng.normal(scale=0.6, size=(6, 4)),
                tp=rng.uniform(0.8, 2.0, 5), bp=rng.uniform(0.8, 2.0, 6), st=0.95, sb=1.2, c=1.3,
                res=(10, 11, 9), center=(0.0, 0.05, 0.1), lengths=(2.6, 2.6, 2.6))
    case["weight"] = rng.uniform(0.5, 2.0, 10 * 11 * 9)
    r = oracle_ap(case)
    assert r["overflowed"] == 0
    h = 1e-6
    fdb = []
    for j in range(6):
        s = []
        for sg in (1, -1):
            bp = case["bp"].copy()
            bp[j] += sg * h
            s.append(oracle_ap(case, bp=bp)["score"])
        fdb.append((s[0] - s[1]) / (2 * h))
    assert _rel(r["grad_bottom"], np.array(fdb)) < 1e-6
    s = [oracle_ap(case, sb=case["sb"] + sg * h)["score"] for sg in (1, -1)]
    fd_sb = (s[0] - s[1]) / (2 * h)
    assert abs(r["grad_bottom_scalar"] - fd_sb) < 1e-6 * abs(fd_sb)


def test_swapped_factor_order_would_fail_the_difference_check():
    """Planted mistakes, in a copy of the matrix restatement: the bottom term
    with its prefix and suffix products swapped, and the bulk term with 1/s in
    place of 1/p_j, both move away from central differences of the score."""
    rng = np.random.default_rng(9)
    case = dict(top=rng.normal(scale=0.6, size=(4, 4)), bot=rng.normal(scale=0.6, size=(3, 4)),
                tp=rng.uniform(0.8, 2.0, 4), bp=rng.uniform(0.8, 2.0, 3), st=1.0, sb=1.2, c=1.0,
                **_single_cell((0.3, -0.2, 0.4)))
    case["weight"] = np.array([1.0])
    c = O.cell_center(O.make_grid(case["res"], case["center"], case["lengths"]), 0, 0, 0)
    q0 = np.array([c[0], c[1], c[2], 0.0])
    case["c"] = _scaling_near_surface(q0, case)
    h = 1e-6
    fdb = []
    for j in range(3):
        s = []
        for sg in (1, -1):
            bp = case["bp"].copy()
            bp[j] += sg * h
            s.append(oracle_ap(case, bp=bp)["score"])
        fdb.append((s[0] - s[1]) / (2 * h))
    fdb = np.array(fdb)
    mf = matrix_form_cell(q0, 1.0, case)
    assert _rel(np.asarray(mf["grad_bottom"]), fdb) < 1e-5
    # the same terms with L and R_suf swapped
    bb, BP = _matrix_terms(q0, case["bot"], case["bp"], case["sb"])
    T = _prod(_matrix_terms(q0, case["top"], case["tp"], case["st"])[1])
    R = T @ np.linalg.inv(_prod(BP))
    F = np.log(np.sqrt(qdot(R, R))) + np.log(case["c"])
    dSdF = -ALPHA * (1 - np.tanh(ALPHA * F) ** 2)
    swapped = []
    for j in range(3):
        L, Rs = _prod(BP[:j]), _prod(BP[j + 1:])
        mid = (1.0 / case["sb"]) * BP[j] @ np.linalg.inv(scipy.linalg.logm(bb[j]))
        swapped.append(dSdF * qdot(R, T @ np.linalg.inv(-1.0 * Rs @ mid @ L)) / qdot(R, R))
    assert _rel(np.asarray(swapped), fdb) > 1e-2
    s = [oracle_ap(case, sb=case["sb"] + sg * h)["score"] for sg in (1, -1)]
    fd_sb = (s[0] - s[1]) / (2 * h)
    assert abs(mf["grad_bottom_scalar"] - fd_sb) < 1e-5 * abs(fd_sb)
    assert abs(float(np.sum(fdb)) - fd_sb) > 1e-2 * abs(fd_sb)
