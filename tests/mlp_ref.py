"""Numpy reference of the 12-64-64-1 SiLU MLP (test infrastructure).

forward/backward in float64, optionally rounding the operands of every matrix
product to bf16 exactly where the MFMA kernels do, so the GPU gradient can be
compared tightly; the float64 path is pinned by central finite differences.
"""
import numpy as np

NP, NF, H = 5057, 12, 64


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[0:768].reshape(64, 12), P[768:832], P[832:4928].reshape(64, 64), P[4928:4992],
            P[4992:5056], P[5056])


def silu(z):
    return z / (1 + np.exp(-z))


def dsilu(z):
    s = 1 / (1 + np.exp(-z))
    return s * (1 + z * (1 - s))


def forward(P, norm, X, emulate=False):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    xn = (np.asarray(X, np.float64) - norm[:NF]) / norm[NF:]
    if emulate:
        xn = (np.asarray(X, np.float32) - np.float32(norm[:NF])) * (np.float32(1) / np.float32(norm[NF:]))
    xn = r(xn)
    z1 = xn @ r(W1).T + b1
    h1 = r(silu(z1))
    z2 = h1 @ r(W2).T + b2
    y = silu(z2) @ w3 + b3
    return y, (xn, z1, h1, z2)


LOG2E, LN2 = 1.4426950408889634, 0.6931471805599453


def forward_infer(P, norm, X, rounding=True):
    """The inference kernel's base-2 form with its bf16 operands: z' = z log2 e
    from W1 log2 e and b1 log2 e, h' = z' / (1 + 2^-z') = SiLU(z) / ln 2,
    z2' = W2 h' + b2 log2 e, y = (w3 ln 2) . h2' + b3 -- in exact arithmetic
    (rounding=False: float64 throughout) the same function as forward()."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    if rounding:
        r, f = bf16, np.float32
        xn = (np.asarray(X, np.float32) - np.float32(norm[:NF])) * (np.float32(1) / np.float32(norm[NF:]))
    else:
        r, f = (lambda a: np.asarray(a, dtype=np.float64)), np.float64
        xn = (np.asarray(X, np.float64) - norm[:NF]) / norm[NF:]
    xn = r(xn)
    silu2 = lambda z: z / (1 + np.exp2(-z))  # noqa: E731
    z1 = xn @ r(f(LOG2E) * W1.astype(f)).T + f(LOG2E) * b1.astype(f)
    h1 = r(silu2(z1))
    z2 = h1 @ r(W2).T + f(LOG2E) * b2.astype(f)
    return silu2(z2) @ (f(LN2) * w3.astype(f)) + b3


def loss_grad(P, norm, X, T, emulate=False):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (xn, z1, h1, z2) = forward(P, norm, X, emulate)
    n = len(X)
    e = y - T
    dy = 2 * e / n
    h2 = silu(z2)
    g = np.zeros(NP)
    g[4992:5056] = h2.T @ dy
    g[5056] = dy.sum()
    dz2 = dy[:, None] * w3[None, :] * dsilu(z2)
    g[4928:4992] = dz2.sum(0)
    dz2r = r(dz2)
    g[832:4928] = (dz2r.T @ h1).reshape(-1)
    dh1 = dz2r @ r(W2)
    dz1 = dh1 * dsilu(z1)
    g[768:832] = dz1.sum(0)
    g[0:768] = (r(dz1).T @ xn).reshape(-1)
    return float(np.mean(e * e)), g


def init_params(seed=5):
    g = np.random.default_rng(seed)
    P = np.zeros(NP)
    P[0:768] = g.normal(scale=np.sqrt(2.0 / 12), size=768)
    P[832:4928] = g.normal(scale=np.sqrt(2.0 / 64), size=4096)
    P[4992:5056] = g.normal(scale=np.sqrt(1.0 / 64), size=64)
    return P.astype(np.float32)
