"""Training pipeline and weight files of the learned opening criterion.

The network (12-64-64-1, hard-swish h(z) = z min(max(z/6 + 1/2, 0), 1))
predicts log10 |exact - far field| of a Barnes–Hut node pair from 12
dimensionless features; samples and labels come from the engine itself
(flk_mlp_samples: exact FP64 direct sum over the two nodes' segments against
the FP32 second-order expansion), training runs on the GPU with the engine's
bf16-MFMA forward/backward kernels and Adam.

Weight file: b"FLKMLP2\\0", 5,057 float32 parameters, 24 float32 feature
normalisation values (mean[12], std[12]), little endian.  (b"FLKMLP1\\0"
files hold weights of an earlier SiLU network and are refused.)
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import MLP_NFEAT, MLP_NPARAM, Engine
from . import workloads

MAGIC = b"FLKMLP2\0"
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mlp_criterion.bin")


def save_weights(path: str, params: np.ndarray, norm: np.ndarray) -> None:
    params = np.asarray(params, dtype="<f4").reshape(-1)
    norm = np.asarray(norm, dtype="<f4").reshape(-1)
    assert params.size == MLP_NPARAM and norm.size == 2 * MLP_NFEAT
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(params.tobytes())
        f.write(norm.tobytes())


def load_weights(path: str = DEFAULT_WEIGHTS) -> tuple[np.ndarray, np.ndarray]:
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] == b"FLKMLP1\0":
        raise ValueError(f"{path}: weights of the earlier SiLU network; retrain "
                         "(python -m fastlinkingnumbers_amd.train_criterion)")
    if data[:8] != MAGIC or len(data) != 8 + 4 * (MLP_NPARAM + 2 * MLP_NFEAT):
        raise ValueError(f"{path}: not an MLP weight file")
    arr = np.frombuffer(data[8:], dtype="<f4")
    return arr[:MLP_NPARAM].copy(), arr[MLP_NPARAM:].copy()


def init_params(seed: int = 5) -> np.ndarray:
    """He-style initialisation of the 5,057 parameters (biases zero)."""
    g = np.random.default_rng(seed)
    P = np.zeros(MLP_NPARAM, dtype=np.float32)
    P[0:768] = g.normal(scale=np.sqrt(2.0 / 12), size=768)
    P[832:4928] = g.normal(scale=np.sqrt(2.0 / 64), size=4096)
    P[4992:5056] = g.normal(scale=np.sqrt(1.0 / 64), size=64)
    return P


def make_samples(eng: Engine, n_pairs: int = 6, n_vertices: int = 200_000, per_pair: int = 400_000,
                 seed: int = 5, rg: float = 100.0, log=print) -> tuple[np.ndarray, np.ndarray]:
    """Node-pair samples from workload-3-style random-walk pairs."""
    xs, ys = [], []
    for k in range(n_pairs):
        curves = workloads.random_walk_pair(seed=1000 + seed * 97 + k, n=n_vertices, rg=rg, workload=5)
        eng.set_curves(curves)
        t = time.time()
        x, y = eng.mlp_samples(0, 1, per_pair, seed=seed * 131 + k)
        log(f"pair {k}: {len(x)} samples in {time.time() - t:.1f}s")
        xs.append(x)
        ys.append(y)
    return np.concatenate(xs), np.concatenate(ys)


def split(x, y, frac_train=0.8, seed=5):
    g = np.random.default_rng(seed)
    idx = g.permutation(len(x))
    k = int(frac_train * len(x))
    return x[idx[:k]], y[idx[:k]], x[idx[k:]], y[idx[k:]]


def train(eng: Engine, x, y, epochs=30, batch=8192, lr=2e-3, seed=5, log=print):
    """Trains from scratch on (x, y); returns (params, norm, train_loss)."""
    mean = x.mean(axis=0)
    std = x.std(axis=0) + 1e-6
    norm = np.concatenate([mean, std]).astype(np.float32)
    eng.mlp_set_weights(init_params(seed), norm)
    loss = None
    for stage, (ep, rate) in enumerate([(epochs // 2, lr), (epochs - epochs // 2, lr / 5)]):
        t = time.time()
        loss = eng.mlp_train(x, y, epochs=ep, batch=batch, lr=rate, seed=seed + stage)
        log(f"stage {stage}: {ep} epochs lr {rate:g}: train mse {loss:.4f} ({time.time() - t:.1f}s)")
    params, norm = eng.mlp_get_weights()
    return params, norm, loss
