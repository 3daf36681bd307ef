"""Direct parity of the device Barnes–Hut trees, moments, far field and
traversal (reference src/compute_link.cpp:294-389, :410-524, :695-841).

flk_tree_moments returns the device's own tree records.  Every node's moments
must equal a brute-force integration of its segments about its centre (FP32
record tolerance), and the host restatement of the traversal over those
records (tests/bh_host.py: the oracle's FP64 far field and leaf term on the
device's node pairs; it reproduces the oracle's own Barnes–Hut exactly on the
oracle's trees) must reproduce the device's linking integral, its Eq. 14
estimate and its visit, far-field and leaf counts.  A dropped or sign-flipped
dipole or quadrupole term moves the device result by that term's size, far
above the FP32 tolerance used here.
"""
import numpy as np
import pytest

import bh_host as H

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

_GL = (0.5 - 0.5 / np.sqrt(3.0), 0.5 + 0.5 / np.sqrt(3.0))


def _brute_moments(pts, segs, ctr):
    """c = int r', D_kj = int r'_k (r - ctr)_j, Q_kij = int r'_k (r-ctr)_i (r-ctr)_j
    over the segments; two-point Gauss-Legendre is exact for these
    polynomials."""
    a = pts[segs]
    t = pts[(segs + 1) % len(pts)] - a
    c = t.sum(0)
    D = np.zeros((3, 3))
    Q = np.zeros((3, 3, 3))
    for u in _GL:
        r = a + u * t - ctr
        D += 0.5 * np.einsum("sk,sj->kj", t, r)
        Q += 0.5 * np.einsum("sk,si,sj->kij", t, r, r)
    return c, D, Q, np.abs(t).sum()


def _curves(shift):
    c, d = W.random_walk_pair(seed=7, n=1500, rg=12.0, workload=3)
    return [c + shift, d + shift]


@pytest.mark.parametrize("leaf", [1, 3])
@pytest.mark.parametrize("shift", [0.0, 4096.0])
def test_device_tree_records_match_brute_force(engine, leaf, shift):
    """Structure (contiguous median-split ranges, leaf buckets, a permutation
    of the segments) and every node's centre, radius, c, D, Q and norms.
    The shifted copy sits far from the origin, where the engine stores
    coordinates relative to a snapped origin."""
    curves = _curves(shift)
    engine.set_curves(curves)
    pts = curves[0]
    n = len(pts)
    nodes, order = engine.tree_moments(0, leaf_size=leaf)
    assert sorted(order.tolist()) == list(range(n))
    N = len(nodes["radius"])
    assert nodes["seg0"][0] == 0 and nodes["nseg"][0] == n
    seen_leaf = 0
    worst = {"c": 0.0, "D": 0.0, "Q": 0.0}
    for k in range(N):
        s0, ns = int(nodes["seg0"][k]), int(nodes["nseg"][k])
        segs = order[s0:s0 + ns].astype(np.int64)
        ch = nodes["child"][k]
        if ch[0] < 0:
            assert 1 <= ns <= leaf
            seen_leaf += ns
        else:
            assert ns > leaf
            l, r = int(ch[0]), int(ch[1])
            assert nodes["seg0"][l] == s0 and nodes["nseg"][l] == ns // 2
            assert nodes["seg0"][r] == s0 + ns // 2 and nodes["nseg"][r] == ns - ns // 2
        ends = np.concatenate([pts[segs], pts[(segs + 1) % n]])
        lo, hi = ends.min(0), ends.max(0)
        np.testing.assert_allclose(nodes["center"][k], 0.5 * (lo + hi), rtol=0, atol=1e-12 * (1 + shift))
        rad = 0.5 * np.linalg.norm(hi - lo)
        assert abs(nodes["radius"][k] - rad) <= 2e-7 * rad
        c, D, Q, sc = _brute_moments(pts, segs, nodes["center"][k])
        R = max(rad, 1e-30)
        # FP32 records: errors relative to the node's own scale (sum |t|,
        # times the radius per order)
        for key, got, ref, scale in (("c", nodes["c"][k], c, sc), ("D", nodes["D"][k], D, sc * R),
                                     ("Q", nodes["Q"][k], Q, sc * R * R)):
            e = float(np.max(np.abs(got - ref))) / scale
            worst[key] = max(worst[key], e)
        assert abs(nodes["cn"][k] - np.linalg.norm(c)) <= 1e-5 * sc
        assert abs(nodes["dn"][k] - np.linalg.norm(D) / np.sqrt(3)) <= 1e-5 * sc * R
        assert abs(nodes["qn"][k] - np.sqrt(np.sum(Q * Q)) / 3) <= 1e-5 * sc * R * R
    assert seen_leaf == n
    print(f"leaf {leaf} shift {shift}: {N} nodes, worst relative moment errors {worst}")
    assert worst["c"] < 1e-6 and worst["D"] < 2e-6 and worst["Q"] < 4e-6


@pytest.mark.parametrize("leaf,beta", [(1, 2.0), (4, 2.0), (1, 1.5)])
def test_device_traversal_matches_host_restatement(engine, leaf, beta):
    """The device linking integral, error estimate and counts equal the host
    restatement's over the device's exported trees (beta fixed: no retry)."""
    curves = _curves(0.0)
    engine.set_curves(curves)
    T1 = H.from_device(*engine.tree_moments(0, leaf_size=leaf), curves[0])
    T2 = H.from_device(*engine.tree_moments(1, leaf_size=leaf), curves[1])
    host = H.traverse(T1, T2, beta, fp32_test=True)
    res = engine.linking_numbers([[0, 1]], method="barneshut", leaf_size=leaf, beta_init=beta, beta_limit=beta,
                                 fallback_tol=0.0)
    s = res.stats
    print(f"leaf {leaf} beta {beta}: device {res.lk[0]:.9f} host {host['lk']:.9f} "
          f"(far-field scale {host['far_abs']:.3g}); visits {s['bh_visits']} / {host['visits']}, "
          f"far {s['bh_far']} / {host['far']}, leaf pairs {s['bh_leaf_pairs']} / {host['leaf_pairs']}; "
          f"est {res.err[0]:.6g} / {0.05 * host['est']:.6g}")
    assert s["retries"] == 0
    assert s["bh_visits"] == host["visits"] and s["bh_far"] == host["far"]
    assert s["bh_leaf_pairs"] == host["leaf_pairs"]
    # FP32 terms: a few units of FP32 rounding of the summed magnitudes
    assert abs(res.lk[0] - host["lk"]) < 1e-6 * (1.0 + host["far_abs"])
    # the pair's estimate is the summed Eq. 14 bound times err_scale 0.05
    assert abs(res.err[0] - 0.05 * host["est"]) <= 1e-5 * 0.05 * host["est"]
