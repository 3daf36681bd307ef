"""Seeded synthetic curve collections for the five benchmark workloads.

Every generator uses numpy's PCG64 stream seeded by the workload number (plus
a per-call seed), so the same call gives the same curves on any machine.
Curves are closed polylines returned as (n, 3) float64 arrays; the fp32
workloads are rounded to float32 first (their coordinates are exactly
representable in FP32, as if read from a BCC file).

1. hopf_torus_circles   10 loops, fp64: Hopf pair (LK +-1), (2,4) torus link
                        (LK +-2), 6 unlinked random circles.
2. chainmail            40 x 40 sheet of wobbly rings, pitch 1.5, rings tilted
                        +-45 deg (checkerboard) about the in-plane diagonal, so
                        each interior ring links its 4 grid neighbours.
3. random_walk_pair     2 closed Gaussian random walks (drift-corrected),
                        rescaled to a radius of gyration, interpenetrating.
4. helical_lattice      rings with 3 random Fourier modes on a helical
                        lattice (the chain-mail sheet wrapped into a tube).
5. random_walk_pair     as 3 at 8M vertices and radius of gyration 200, plus
                        smaller training pairs for the learned criterion.
"""
from __future__ import annotations

import numpy as np


def _rng(workload: int, seed: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.PCG64([workload, seed]))


def _rot(axis, ang) -> np.ndarray:
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
    return np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * (K @ K)


def _f32(c: np.ndarray) -> np.ndarray:
    return c.astype(np.float32).astype(np.float64)


def _circle(n: int) -> np.ndarray:
    t = 2 * np.pi * np.arange(n) / n
    return np.stack([np.cos(t), np.sin(t), np.zeros(n)], 1)


# ---------------------------------------------------------------- workload 1
def hopf_torus_circles(seed: int = 0, n_circle: int = 2000, n_torus: int = 4000,
                       jitter: float = 1e-3) -> list[np.ndarray]:
    g = _rng(1, seed)
    curves = []
    c0 = _circle(n_circle)                       # z = 0 plane, centre 0
    c1 = _circle(n_circle)[:, [0, 2, 1]] + [1.0, 0.0, 0.0]  # y = 0 plane, centre (1,0,0)
    curves += [c0, c1]
    # (2,4) torus link: two (1,2) curves on a torus R=2, r=0.5, half a turn apart
    t = 2 * np.pi * np.arange(n_torus) / n_torus
    for m in range(2):
        th = 2 * t + m * np.pi
        curves.append(np.stack([(2 + 0.5 * np.cos(th)) * np.cos(t),
                                (2 + 0.5 * np.cos(th)) * np.sin(t),
                                0.5 * np.sin(th)], 1) + [-6.0, 0.0, 0.0])
    # 6 unit circles, random normals, centres U([0,20]^3); a circle is kept only
    # when its 1.1-ball holds no point of an earlier curve (clearance >= 0.1,
    # and no earlier curve can pass through its disk, so it is unlinked).
    pts = np.concatenate(curves)
    while len(curves) < 10:
        v = g.normal(size=3)
        v /= np.linalg.norm(v)
        ctr = g.uniform(0, 20, size=3)
        if np.min(np.linalg.norm(pts - ctr, axis=1)) < 1.1:
            continue
        a = np.cross(v, [1.0, 0, 0] if abs(v[0]) < 0.9 else [0, 1.0, 0])
        a /= np.linalg.norm(a)
        b = np.cross(v, a)
        tt = 2 * np.pi * np.arange(n_circle) / n_circle
        c = ctr + np.outer(np.cos(tt), a) + np.outer(np.sin(tt), b)
        curves.append(c)
        pts = np.concatenate([pts, c])
    return [c + g.normal(scale=jitter, size=c.shape) for c in curves]


# ---------------------------------------------------------------- workload 2
def _loglog_counts(g, count, nmin, nmax):
    return np.rint(np.exp(g.uniform(np.log(nmin), np.log(nmax), size=count))).astype(np.int64)


def chainmail(seed: int = 0, grid: int = 40, nmin: int = 256, nmax: int = 65536,
              pitch: float = 1.5, amp: float = 0.05) -> list[np.ndarray]:
    g = _rng(2, seed)
    ns = _loglog_counts(g, grid * grid, nmin, nmax)
    ks = g.integers(3, 8, size=grid * grid)
    phis = g.uniform(0, 2 * np.pi, size=grid * grid)
    Rp, Rm = _rot((1, 1, 0), np.pi / 4), _rot((1, 1, 0), -np.pi / 4)
    out = []
    for i in range(grid):
        for j in range(grid):
            k = i * grid + j
            t = 2 * np.pi * np.arange(ns[k]) / ns[k]
            p = np.stack([np.cos(t), np.sin(t), amp * np.sin(ks[k] * t + phis[k])], 1)
            R = Rp if (i + j) % 2 == 0 else Rm
            out.append(_f32(p @ R.T + [pitch * i, pitch * j, 0.0]))
    return out


# ---------------------------------------------------------------- workloads 3, 5
def random_walk(g: np.random.Generator, n: int, rg: float) -> np.ndarray:
    steps = g.normal(size=(n, 3))
    x = np.concatenate([[np.zeros(3)], np.cumsum(steps, axis=0)])  # x_0 .. x_n
    k = np.arange(n + 1)[:, None] / n
    x = x - k * x[-1]                                          # closes: x_n = x_0
    x = x[:-1]
    x -= x.mean(axis=0)
    x *= rg / np.sqrt(np.mean(np.sum(x * x, axis=1)))
    return x


def random_walk_pair(seed: int = 0, n: int = 2_000_000, rg: float = 100.0, workload: int = 3,
                     offset: float = 0.25) -> list[np.ndarray]:
    g = _rng(workload, seed)
    a = random_walk(g, n, rg)
    b = random_walk(g, n, rg)
    d = g.normal(size=3)
    d *= offset * rg / np.linalg.norm(d)
    return [_f32(a), _f32(b + d)]


# ---------------------------------------------------------------- workload 4
def helical_lattice(seed: int = 0, nloops: int = 10000, per_turn: int = 101, nmin: int = 64,
                    nmax: int = 262144, pitch: float = 1.6, amp: float = 0.1,
                    only=None) -> list[np.ndarray]:
    """The chain-mail sheet wrapped into a tube: ring k sits at angle
    2 pi k / per_turn on a cylinder whose circumference is per_turn * pitch,
    rising pitch per turn, so ring k neighbours k +- 1 along the helix and
    k +- per_turn across turns (per_turn odd keeps the tilt alternating both
    ways).  Each ring carries 3 random Fourier modes (total amplitude <= amp).
    At pitch 1.6 a ring links 0-4 of its neighbours (mostly 2-3, LK +-1): the
    Fourier modes unlink some neighbour pairs that pitch 1.5 keeps linked.
    only: ring indices to build (in that order), the others' random draws
    still consumed, so each ring is the one of the whole lattice."""
    g = _rng(4, seed)
    ns = _loglog_counts(g, nloops, nmin, nmax)
    Rt = per_turn * pitch / (2 * np.pi)
    Rp, Rm = _rot((1, 1, 0), np.pi / 4), _rot((1, 1, 0), -np.pi / 4)
    want = None if only is None else set(int(k) for k in only)
    out, built = [], {}
    for k in range(nloops):
        n = ns[k]
        freq = g.integers(2, 7, size=(3, 2))
        ph = g.uniform(0, 2 * np.pi, size=(3, 2))
        am = g.uniform(0, amp / 3, size=(3, 2))
        if want is not None and k not in want:
            continue
        t = 2 * np.pi * np.arange(n) / n
        r = 1 + sum(am[m, 0] * np.cos(freq[m, 0] * t + ph[m, 0]) for m in range(3))
        w = sum(am[m, 1] * np.sin(freq[m, 1] * t + ph[m, 1]) for m in range(3))
        p = np.stack([r * np.cos(t), r * np.sin(t), w], 1)
        p = p @ (Rp if k % 2 == 0 else Rm).T     # sheet frame: x along helix, y up, z out
        th = 2 * np.pi * k / per_turn
        z = pitch * k / per_turn
        e_t = np.array([-np.sin(th), np.cos(th), 0.0])
        e_z = np.array([0.0, 0.0, 1.0])
        e_r = np.array([np.cos(th), np.sin(th), 0.0])
        ctr = Rt * e_r + z * e_z
        ring = _f32(ctr + np.outer(p[:, 0], e_t) + np.outer(p[:, 1], e_z) + np.outer(p[:, 2], e_r))
        if want is None:
            out.append(ring)
        else:
            built[k] = ring
    return out if want is None else [built[int(k)] for k in only]


def workload(num: int, scale: float = 1.0, seed: int = 0) -> list[np.ndarray]:
    """Workload `num` (1-5) at full size (scale=1) or with vertex counts and
    collection sizes scaled down for tests."""
    if num == 1:
        return hopf_torus_circles(seed, n_circle=max(16, int(2000 * scale)),
                                  n_torus=max(32, int(4000 * scale)))
    if num == 2:
        grid = max(2, int(round(40 * min(1.0, scale ** 0.5))))
        return chainmail(seed, grid=grid, nmin=max(32, int(256 * scale)),
                         nmax=max(64, int(65536 * scale)))
    if num == 3:
        return random_walk_pair(seed, n=max(64, int(2_000_000 * scale)), rg=100.0, workload=3)
    if num == 4:
        return helical_lattice(seed, nloops=max(16, int(10000 * scale)),
                               nmin=max(16, int(64 * scale)), nmax=max(64, int(262144 * scale)))
    if num == 5:
        return random_walk_pair(seed, n=max(64, int(8_000_000 * scale)), rg=200.0, workload=5)
    raise ValueError(num)
