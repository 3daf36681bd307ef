"""Trains the learned opening criterion on the GPU and writes its weight file.

    python -m fastlinkingnumbers_amd.train_criterion --out PATH

Samples: node pairs of workload-3-style random-walk loop pairs labelled by the
engine (exact FP64 sums vs FP32 far field); 80/20 seeded split; training with
the engine's bf16-MFMA kernels.  Prints the train/validation errors.
"""
import argparse
import json
import time

import numpy as np

from . import Engine
from . import criterion


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=criterion.DEFAULT_WEIGHTS)
    ap.add_argument("--pairs", type=int, default=10)
    ap.add_argument("--vertices", type=int, default=400_000)
    ap.add_argument("--per-pair", type=int, default=500_000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--seed", type=int, default=5)
    a = ap.parse_args()
    eng = Engine(0)
    t0 = time.time()
    x, y = criterion.make_samples(eng, a.pairs, a.vertices, a.per_pair, seed=a.seed)
    t1 = time.time()
    xt, yt, xv, yv = criterion.split(x, y, 0.8, seed=a.seed)
    params, norm, loss = criterion.train(eng, xt, yt, epochs=a.epochs, seed=a.seed)
    t2 = time.time()
    pv = eng.mlp_forward(xv)
    val = float(np.mean((pv - yv) ** 2))
    base = float(np.mean((yv - yt.mean()) ** 2))
    criterion.save_weights(a.out, params, norm)
    print(json.dumps({"samples": int(len(x)), "train": int(len(xt)), "validation": int(len(xv)),
                      "train_mse": loss, "val_mse": val, "val_baseline_mse": base,
                      "val_mae_log10": float(np.mean(np.abs(pv - yv))),
                      "sample_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1), "out": a.out}))
    eng.close()


if __name__ == "__main__":
    main()
