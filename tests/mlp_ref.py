"""Numpy reference of the 12-64-64-1 hard-swish MLP (test infrastructure).

forward/backward in float64, optionally rounding the operands of every matrix
product to bf16 exactly where the MFMA training kernels do, so the GPU
gradient can be compared tightly; the float64 path is pinned by central
finite differences.  forward_infer models the inference kernels' FP16 form.
"""
import numpy as np

NP, NF, H = 5057, 12, 64


def bf16(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = a.view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def f16(a):
    """Round to IEEE binary16, nearest even (numpy converts float64 directly)."""
    return np.asarray(a, dtype=np.float64).astype(np.float16).astype(np.float64)


def unpack(P):
    P = np.asarray(P, dtype=np.float64)
    return (P[0:768].reshape(64, 12), P[768:832], P[832:4928].reshape(64, 64), P[4928:4992],
            P[4992:5056], P[5056])


def hswish(z):
    return z * np.clip(z / 6 + 0.5, 0.0, 1.0)


def dhswish(z):
    return np.where(z <= -3, 0.0, np.where(z < 3, z / 3 + 0.5, 1.0))


def forward(P, norm, X, emulate=False):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    xn = (np.asarray(X, np.float64) - norm[:NF]) / norm[NF:]
    if emulate:
        xn = (np.asarray(X, np.float32) - np.float32(norm[:NF])) * (np.float32(1) / np.float32(norm[NF:]))
    xn = r(xn)
    z1 = xn @ r(W1).T + b1
    h1 = r(hswish(z1))
    z2 = h1 @ r(W2).T + b2
    y = hswish(z2) @ w3 + b3
    return y, (xn, z1, h1, z2)


SIXTH16 = float(np.float16(1.0 / 6.0))


def hswish16(z):
    """The inference kernels' activation: the FP32 layer output rounded to
    FP16, gate = clamp(fma(z, fp16(1/6), 1/2), 0, 1) rounded once, product
    rounded once (products of FP16 numbers are exact in float64)."""
    z = f16(z)
    gate = np.clip(f16(z * SIXTH16 + 0.5), 0.0, 1.0)
    return f16(z * gate)


def forward_infer(P, norm, X):
    """The inference kernels' FP16 form: normalised inputs, weights and
    activations rounded to FP16, FP32 biases and accumulation (float64 here)."""
    W1, b1, W2, b2, w3, b3 = unpack(P)
    xn = (np.asarray(X, np.float32) - np.float32(norm[:NF])) * (np.float32(1) / np.float32(norm[NF:]))
    xn = f16(xn)
    z1 = xn @ f16(W1).T + b1.astype(np.float32)
    h1 = hswish16(z1)
    z2 = h1 @ f16(W2).T + b2.astype(np.float32)
    return hswish16(z2) @ f16(w3) + np.float32(b3)


def loss_grad(P, norm, X, T, emulate=False):
    W1, b1, W2, b2, w3, b3 = unpack(P)
    r = bf16 if emulate else (lambda a: np.asarray(a, dtype=np.float64))
    y, (xn, z1, h1, z2) = forward(P, norm, X, emulate)
    n = len(X)
    e = y - T
    dy = 2 * e / n
    h2 = hswish(z2)
    g = np.zeros(NP)
    g[4992:5056] = h2.T @ dy
    g[5056] = dy.sum()
    dz2 = dy[:, None] * w3[None, :] * dhswish(z2)
    g[4928:4992] = dz2.sum(0)
    dz2r = r(dz2)
    g[832:4928] = (dz2r.T @ h1).reshape(-1)
    dh1 = dz2r @ r(W2)
    dz1 = dh1 * dhswish(z1)
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
