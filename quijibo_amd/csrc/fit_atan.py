"""Coefficients of the FP32 polynomial in qj_fp32.h::angle_upper.

atan(t) = t P(t^2) on [0, 1], P of degree 8, fitted by least squares to
atan(t)/t with relative weight at n Chebyshev nodes of s = t^2 in [0, 1].
Run as a script to print the coefficients (lowest degree first) and the
largest error of the FP32 Horner evaluation against float64 atan.
"""
import numpy as np


def coefficients(deg=8, n=4000):
    k = np.arange(n)
    s = 0.5 + 0.5 * np.cos((2 * k + 1) * np.pi / (2 * n))
    t = np.sqrt(s)
    f = np.arctan(t) / t
    return np.polynomial.polynomial.polyfit(s, f, deg, w=1.0 / f).astype(np.float32)


def horner32(c, t):
    t = np.asarray(t, dtype=np.float32)
    s = (t * t).astype(np.float32)
    p = np.full_like(t, c[-1])
    for ci in c[-2::-1]:
        p = (p * s + ci).astype(np.float32)
    return (t * p).astype(np.float32)


def max_ulp(c, n=2000001):
    t = np.linspace(0.0, 1.0, n).astype(np.float32)[1:]
    ref = np.arctan(t.astype(np.float64))
    return float(np.max(np.abs(horner32(c, t) - ref) / np.spacing(ref.astype(np.float32)).astype(np.float64)))


if __name__ == "__main__":
    c = coefficients()
    print([float(x) for x in c])
    print("max ulp %.2f" % max_ulp(c))
