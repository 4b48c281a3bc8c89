"""Pins the oracle's quaternion algebra (QUATERNION.cpp, POLYNOMIAL_4D.cpp) with
closed forms and an independent representation: a quaternion a+bi+cj+dk is the
2x2 complex matrix [[a+bi, c+di], [-c+di, a-bi]], whose matrix product is the
Hamilton product and whose principal matrix power is the polar power."""
import numpy as np
import pytest
import scipy.linalg

import oracle as O


def M(q):
    a, b, c, d = q
    return np.array([[a + 1j * b, c + 1j * d], [-c + 1j * d, a - 1j * b]])


def unM(m):
    return np.array([m[0, 0].real, m[0, 0].imag, m[0, 1].real, m[0, 1].imag])


rng = np.random.default_rng(7)


def rq(n=None):
    return rng.normal(size=(4,) if n is None else (n, 4))


def test_basis_products():
    one, i, j, k = np.eye(4)
    assert np.array_equal(O.qmul(i, j), k)
    assert np.array_equal(O.qmul(j, k), i)
    assert np.array_equal(O.qmul(k, i), j)
    assert np.array_equal(O.qmul(j, i), -k)
    assert np.array_equal(O.qmul(i, i), -one)


def test_product_matches_matrix_representation():
    for _ in range(50):
        p, q = rq(), rq()
        assert np.allclose(O.qmul(p, q), unM(M(p) @ M(q)), rtol=1e-14, atol=1e-14)
        assert not np.allclose(O.qmul(p, q), O.qmul(q, p))  # non-commutative


def test_division_identities():
    for _ in range(50):
        p, q = rq(), rq()
        assert np.allclose(O.qdiv(q, q), [1, 0, 0, 0], atol=1e-14)
        assert np.allclose(O.qmul(O.qdiv(p, q), q), p, rtol=1e-13, atol=1e-13)  # right division
        assert np.isclose(O.qmag(O.qdiv(p, q)), O.qmag(p) / O.qmag(q), rtol=1e-14)
        assert np.allclose(O.qdiv(p, q), unM(M(p) @ np.linalg.inv(M(q))), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("e", [1.0, 2.0, 3.0, 0.5, 1.7, 2.25])
def test_pow_is_principal_power(e):
    for _ in range(30):
        q = rq()
        ref = unM(scipy.linalg.fractional_matrix_power(M(q), e))
        assert np.allclose(O.qpow(q, e), ref, rtol=1e-11, atol=1e-11)
        assert np.isclose(O.qmag(O.qpow(q, e)), O.qmag(q) ** e, rtol=1e-13)


def test_pow_integer_and_additive():
    for _ in range(30):
        q = rq()
        assert np.allclose(O.qpow(q, 1.0), q, rtol=1e-13, atol=1e-14)
        assert np.allclose(O.qpow(q, 2.0), O.qmul(q, q), rtol=1e-13, atol=1e-13)
        assert np.allclose(O.qmul(O.qpow(q, 0.7), O.qpow(q, 1.6)), O.qpow(q, 2.3), rtol=1e-12, atol=1e-12)


def test_pow_reference_guards():
    # purely real negative base: the reference returns +|q|^e (QUATERNION.cpp:381-391)
    p = O.qpow([-2.0, 0, 0, 0], 3.0)
    assert np.isclose(p[0], 8.0, rtol=1e-15) and np.array_equal(p[1:], [0.0, 0.0, 0.0])
    # zero base: acos(0/0) -> NaN
    assert np.isnan(O.qpow([0.0, 0, 0, 0], 2.0)).any()
    # negative exponent: magnitude <= 0 -> imaginary part dropped
    q = np.array([0.3, 0.4, -0.2, 0.1])
    p = O.qpow(q, -1.5)
    assert np.array_equal(p[1:], [0.0, 0.0, 0.0])
    th = np.arccos(q[0] / np.linalg.norm(q))
    assert np.isclose(p[0], np.linalg.norm(q) ** -1.5 * np.cos(-1.5 * th), rtol=1e-14)


def test_log_inverts_exp_closed_form():
    for _ in range(20):
        q = rq()
        L = O.qlog(q)
        # log q = (log|q|, v/|v| acos(w/|q|)); |v| goes through sqrtf in the reference
        v = q[1:]
        nv = float(np.sqrt(np.float32(v @ v)))
        assert np.isclose(L[0], np.log(np.linalg.norm(q)), rtol=1e-15)
        assert np.allclose(L[1:], v / nv * np.arccos(q[0] / np.linalg.norm(q)), rtol=1e-15)
        ref = unM(scipy.linalg.logm(M(q)))
        assert np.allclose(L, ref, rtol=1e-6, atol=1e-6)


def test_factored_product_order():
    roots = rq(3)
    powers = np.array([1.0, 2.0, 1.5])
    s = 0.8
    p = O.poly(roots, powers, s)
    q = rq()
    terms = [scipy.linalg.fractional_matrix_power(M(q - r), s * e) for r, e in zip(roots, powers)]
    ref = unM(terms[0] @ terms[1] @ terms[2])
    rev = unM(terms[2] @ terms[1] @ terms[0])
    got = O.poly_eval(p, q)
    assert np.allclose(got, ref, rtol=1e-11, atol=1e-11)
    assert not np.allclose(got, rev, rtol=1e-6)


def test_rational_known_answer():
    """POLYNOMIAL_4D::rationalTest (POLYNOMIAL_4D.cpp:520-577): g = z^2 - 1, h = z^2,
    g/h = 1 - 1/z^2 at z = 1.23 + 4.56i, here through the factored path."""
    z = 1.23 + 4.56j
    top = O.poly([[1, 0, 0, 0], [-1, 0, 0, 0]], [1, 1])
    bot = O.poly([[0, 0, 0, 0]], [2])
    q = np.array([z.real, z.imag, 0, 0])
    T = O.poly_eval(top, q)
    B = O.poly_eval(bot, q)
    r = O.qdiv(T, B)
    want = 1 - 1 / z ** 2
    assert np.allclose(r, [want.real, want.imag, 0, 0], rtol=1e-14, atol=1e-15)
