"""Writes tests/golden/gauss_pairs.json: golden linking-number contributions of
random segment pairs from adaptive mpmath quadrature of the Gauss integral

    (1/4 pi) int_0^1 int_0^1 (p(s) - q(t)) . (p' x q') / |p(s) - q(t)|^3 ds dt

(p from p0 to p1, q from q0 to q1).  Independent of the closed form the
engine and the oracle use.  Run: python tests/golden/make_gauss_pairs.py
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 30


def gauss_pair(p0, p1, q0, q1):
    p0, p1, q0, q1 = (mp.matrix(v) for v in (p0, p1, q0, q1))
    dp, dq = p1 - p0, q1 - q0
    cr = mp.matrix([dp[1] * dq[2] - dp[2] * dq[1], dp[2] * dq[0] - dp[0] * dq[2],
                    dp[0] * dq[1] - dp[1] * dq[0]])

    def f(s, t):
        r = (p0 + s * dp) - (q0 + t * dq)
        n = mp.sqrt(r[0] ** 2 + r[1] ** 2 + r[2] ** 2)
        return (r[0] * cr[0] + r[1] * cr[1] + r[2] * cr[2]) / n ** 3

    return mp.quad(f, [0, 1], [0, 1]) / (4 * mp.pi)


def main():
    g = np.random.default_rng(20210806)
    out = []
    for k in range(24):
        p0, q0 = g.normal(size=3), g.normal(size=3)
        scale = [0.2, 1.0, 3.0][k % 3]
        p1 = p0 + scale * g.normal(size=3)
        q1 = q0 + scale * g.normal(size=3)
        v = gauss_pair(p0, p1, q0, q1)
        out.append({"p0": p0.tolist(), "p1": p1.tolist(), "q0": q0.tolist(), "q1": q1.tolist(),
                    "lk": float(v)})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gauss_pairs.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path, len(out))


if __name__ == "__main__":
    main()
