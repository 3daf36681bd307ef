#!/usr/bin/env python3
"""Fits the arctangent polynomials of the direct-sum tile kernel (ds.hip).

atan(t) = t * P(t^2) with P of degree n, minimising the maximum relative
error on |t| <= a (Remez exchange on f(s) = atan(sqrt(s)) / sqrt(s),
s = t^2, weight 1/f), then rounds the coefficients to FP32:

* atan_unit (ds.hip): a = 1, degree 7;
* the paired small-angle series (ds.hip, short path): a = 16/63, degree 4.

Usage: python fit_atan_poly.py  -- prints both coefficient sets (highest
degree first, as ds.hip evaluates them by Horner's rule) and their maximum
relative errors in exact arithmetic.
"""
import numpy as np


def _f(s):
    s = np.asarray(s, dtype=np.float64)
    r = np.sqrt(s)
    out = np.ones_like(s)
    nz = r > 1e-8
    out[nz] = np.arctan(r[nz]) / r[nz]
    small = ~nz  # series 1 - s/3 + s^2/5
    out[small] = 1.0 - s[small] / 3.0 + s[small] ** 2 / 5.0
    return out


def remez(a, degree, iters=60, grid=200001):
    """Coefficients c[0..degree] (lowest first) of P minimising
    max |t P(t^2) / atan(t) - 1| on 0 < t <= a."""
    b = a * a
    m = degree + 2
    # Chebyshev extrema as the first reference
    x = 0.5 * b * (1 - np.cos(np.pi * np.arange(m) / (m - 1)))
    s = np.linspace(0.0, b, grid)
    fs = _f(s)
    c = None
    for _ in range(iters):
        fx = _f(x)
        A = np.zeros((m, m))
        for j in range(degree + 1):
            A[:, j] = x ** j
        A[:, degree + 1] = ((-1.0) ** np.arange(m)) * fx
        sol = np.linalg.solve(A, fx)
        c = sol[: degree + 1]
        err = np.polyval(c[::-1], s) / fs - 1.0
        # new reference: the extremum of each run of constant sign
        sign = np.sign(err)
        sign[sign == 0] = 1
        idx, start = [], 0
        for k in range(1, len(s) + 1):
            if k == len(s) or sign[k] != sign[start]:
                seg = np.arange(start, k)
                idx.append(seg[np.argmax(np.abs(err[seg]))])
                start = k
        if len(idx) < m:
            break
        # keep m alternating extrema with the largest errors
        while len(idx) > m:
            drop = 0 if abs(err[idx[0]]) < abs(err[idx[-1]]) else len(idx) - 1
            idx.pop(drop)
        xn = s[idx]
        if np.allclose(xn, x, rtol=0, atol=1e-15 * b):
            break
        x = xn
    return c


def max_rel_error(coef_high_first, a, n=400001, fp32=False):
    """Maximum relative error of t * P(t^2) against atan(t) on 0 < t <= a,
    in float64 arithmetic or emulating the kernel's FP32 Horner evaluation
    with fused multiply-adds; also the mean signed error (bias)."""
    t = np.linspace(a / n, a, n)
    if fp32:
        t32 = t.astype(np.float32)
        s = (t32 * t32).astype(np.float32)
        p = np.full_like(s, np.float32(coef_high_first[0]))
        for cf in coef_high_first[1:]:
            p = (p.astype(np.float64) * s.astype(np.float64) + np.float64(np.float32(cf))).astype(np.float32)
        v = (t32.astype(np.float64) * p.astype(np.float64)).astype(np.float32).astype(np.float64)
        ref = np.arctan(t32.astype(np.float64))
    else:
        s = t * t
        v = t * np.polyval(np.asarray(coef_high_first, dtype=np.float64), s)
        ref = np.arctan(t)
    rel = v / ref - 1.0
    return float(np.max(np.abs(rel))), float(np.mean(rel))


def main():
    for name, a, deg in (("atan_unit", 1.0, 7), ("small-angle series", 16.0 / 63.0, 4)):
        c = remez(a, deg)
        c32 = [float(np.float32(v)) for v in c[::-1]]
        e64, _ = max_rel_error(list(c[::-1]), a)
        e, bias = max_rel_error(c32, a)
        e32, _ = max_rel_error(c32, a, fp32=True)
        print(f"{name}: |t| <= {a:.6g}, degree {deg} in t^2")
        print("  coefficients (highest degree first):", ", ".join(f"{v:.17g}f" for v in c32))
        print(f"  max relative error: {e64:.3g} (FP64 coefficients), {e:.3g} (FP32 coefficients, mean signed "
              f"{bias:.2g}), {e32:.3g} (evaluated in FP32)")


if __name__ == "__main__":
    main()
