"""Fit the polynomial coefficients of the deterministic FP32 math set.

The library's cone construction and cone-overlap test need acos, asin, atan2,
sin and cos.  Vendor math libraries differ by an ulp here and there, which
would make GPU trees and query results differ from the CPU oracle in rare
borderline cases.  The library therefore evaluates these functions with its own
polynomial approximations built from IEEE-exact operations only (+, -, *, /,
sqrt, comparisons; no fused multiply-add), so that every platform that follows
IEEE 754 single precision produces the same bits.

Method (textbook range reduction + least-squares fit on Chebyshev nodes,
weighted for relative error):
  asin(x) = x + x^3 P(x^2)            on |x| <= 0.5
  asin(x) = pi/2 - 2 asin(sqrt((1-x)/2)) on 0.5 < x <= 1
  atan(x) = x + x^3 Q(x^2)            on |x| <= tan(pi/8), reductions by
            atan(x) = pi/4 + atan((x-1)/(x+1)) and atan(x) = pi/2 - atan(1/x)
  sin(r)  = r + r^3 S(r^2),  cos(r) = 1 - r^2/2 + r^4 C(r^2) on |r| <= pi/4
Run:  python gen_detmath.py   (prints coefficients and float32 error in ulp)
"""
import numpy as np


def cheb_nodes(a, b, n):
    k = np.arange(n)
    x = np.cos((2 * k + 1) * np.pi / (2 * n))
    return 0.5 * (a + b) + 0.5 * (b - a) * x


def series(zs, coef):
    """sum_k coef(k) z^k evaluated in float64 (no cancellation)."""
    out = np.zeros_like(zs)
    for k in range(80):
        out += coef(k) * zs ** k
    return out


def fit(zs, target, deg):
    # least squares in float64, relative weighting handled by caller
    A = np.vstack([zs ** k for k in range(deg + 1)]).T
    c, *_ = np.linalg.lstsq(A, target, rcond=None)
    return c


def horner32(c, z):
    z = np.float32(z)
    acc = np.float32(c[-1])
    for k in range(len(c) - 2, -1, -1):
        acc = np.float32(np.float32(acc * z) + np.float32(c[k]))
    return acc


def ulp_err(approx32, exact64):
    approx32 = approx32.astype(np.float32)
    ulp = np.spacing(np.abs(exact64).astype(np.float32)).astype(np.float64)
    return np.max(np.abs(approx32.astype(np.float64) - exact64) / ulp)


def main():
    n = 4000
    # asin: P(z) with z = x^2 in (0, 0.25]
    zs = cheb_nodes(1e-12, 0.25, n)
    from math import factorial
    tgt = series(zs, lambda k: factorial(2 * k + 2) / (4 ** (k + 1) * factorial(k + 1) ** 2 * (2 * k + 3)))
    P = fit(zs, tgt, 4).astype(np.float32)
    print("asin P:", [float(v) for v in P])
    # atan: Q(z) with z = x^2 in (0, tan(pi/8)^2]
    t8 = np.tan(np.pi / 8)
    zs = cheb_nodes(1e-12, t8 * t8, n)
    tgt = series(zs, lambda k: (-1) ** (k + 1) / (2 * k + 3))
    Q = fit(zs, tgt, 4).astype(np.float32)
    print("atan Q:", [float(v) for v in Q])
    # sin / cos on |r| <= pi/4
    zs = cheb_nodes(1e-12, (np.pi / 4) ** 2, n)
    S = fit(zs, series(zs, lambda k: (-1) ** (k + 1) / factorial(2 * k + 3)), 3).astype(np.float32)
    C = fit(zs, series(zs, lambda k: (-1) ** k / factorial(2 * k + 4)), 3).astype(np.float32)
    print("sin S:", [float(v) for v in S])
    print("cos C:", [float(v) for v in C])

    # float32 evaluation error checks (formula as used by the library)
    x = np.linspace(0, 0.5, 200001).astype(np.float32)
    z = (x * x).astype(np.float32)
    a = (x + (x * z).astype(np.float32) * horner32(P, z)).astype(np.float32)
    print("asin [0,0.5] max ulp:", ulp_err(a[1:], np.arcsin(x[1:].astype(np.float64))))
    x = np.linspace(0, float(t8), 200001).astype(np.float32)
    z = (x * x).astype(np.float32)
    a = (x + (x * z).astype(np.float32) * horner32(Q, z)).astype(np.float32)
    print("atan [0,t8] max ulp:", ulp_err(a[1:], np.arctan(x[1:].astype(np.float64))))
    r = np.linspace(0, np.pi / 4, 200001).astype(np.float32)
    z = (r * r).astype(np.float32)
    s = (r + (r * z).astype(np.float32) * horner32(S, z)).astype(np.float32)
    c = (np.float32(1) - (np.float32(0.5) * z).astype(np.float32) + (z * z).astype(np.float32) * horner32(C, z)).astype(np.float32)
    print("sin max ulp:", ulp_err(s[1:], np.sin(r[1:].astype(np.float64))))
    print("cos max ulp:", ulp_err(c, np.cos(r.astype(np.float64))))


if __name__ == "__main__":
    main()
