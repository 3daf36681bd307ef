"""CPU pins of the oracle (no GPU).

How each oracle function is pinned:
  * end to end (PLS, discretize, DS, BH, adaptive beta, rounding, file format):
    the reference's 5 certificates, byte for byte, and the 5 mean off-integer
    errors test.sh publishes (test.sh:4, :14, :24, :34, :45);
  * per-segment-pair term: independent mpmath quadrature (tests/golden/
    gauss_pairs.json), the rectangle solid-angle closed form, symmetries;
  * whole-curve sums: Hopf +-1, (2,4) torus link +-2, orientation reversal,
    agreement of the two reference direct-sum variants;
  * moments: brute-force integration about each node centre, leaf closed form;
  * far field: an independent numpy second-order Taylor integration of the
    Gauss kernel over the segments, and the convergence order;
  * PLS: brute force vs tree self join, touching boxes;
  * BCC/OBJ import: sizes, fp32 round trip.
"""
import json
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN, ROOT

DATA = os.path.join(GOLDEN, "data")
CERTS = os.path.join(GOLDEN, "reference_certificates")
# test.sh published mean off-integer errors (Barnes-Hut defaults)
MEAN_ERR = {"glove": 0.00214443, "knittubeinit": 0.00366683, "knittubebroken": 0.006343,
            "blenderobj": 0.000668562, "houdiniobj": 0.00138268}


def _input(name):
    if name.endswith("obj"):
        return os.path.join(DATA, name + ".obj"), "PL"
    return os.path.join(DATA, name + ".bcc"), None


@pytest.mark.parametrize("name", list(MEAN_ERR))
def test_fixture_certificate_and_mean_error_bh(tmp_path, name):
    inp, obj = _input(name)
    out = tmp_path / f"{name}.txt"
    r = O.run_cli(inp, str(out), method="barneshut", obj_type=obj)
    with open(os.path.join(CERTS, name + ".txt")) as f:
        assert out.read_text() == f.read()
    line = [l for l in r.stdout.splitlines() if "Mean Off-Integer" in l][0]
    assert line.split(":")[-1].strip() == f"{MEAN_ERR[name]:g}"


@pytest.mark.parametrize("name", ["knittubebroken", "knittubeinit", "blenderobj", "houdiniobj"])
def test_fixture_certificate_ds(tmp_path, name):
    inp, obj = _input(name)
    out = tmp_path / f"{name}.txt"
    O.run_cli(inp, str(out), method="directsum", obj_type=obj)
    with open(os.path.join(CERTS, name + ".txt")) as f:
        assert out.read_text() == f.read()


def test_segment_pair_term_golden_quadrature():
    with open(os.path.join(GOLDEN, "gauss_pairs.json")) as f:
        data = json.load(f)
    for e in data:
        v = O.segment_pair_term(e["p0"], e["p1"], e["q0"], e["q1"])
        assert abs(v - e["lk"]) < 1e-12, e


@pytest.mark.parametrize("a,b,h", [(1.0, 1.0, 1.0), (0.3, 2.0, 0.5), (1.0, 1.0, 1e-3), (5.0, 0.1, 2.0)])
def test_perpendicular_segments_closed_form(a, b, h):
    # solid angle of the 2a x 2b rectangle seen from distance h on its axis
    want = math.asin(a * b / math.sqrt((a * a + h * h) * (b * b + h * h))) / math.pi
    v = O.segment_pair_term([-a, 0, 0], [a, 0, 0], [0, -b, h], [0, b, h])
    assert abs(abs(v) - want) < 1e-13
    # reversing one segment flips the sign, swapping keeps it, mirroring flips it
    assert abs(O.segment_pair_term([a, 0, 0], [-a, 0, 0], [0, -b, h], [0, b, h]) + v) < 1e-13
    assert abs(O.segment_pair_term([0, -b, h], [0, b, h], [-a, 0, 0], [a, 0, 0]) - v) < 1e-13
    assert abs(O.segment_pair_term([-a, 0, 0], [a, 0, 0], [0, -b, -h], [0, b, -h]) + v) < 1e-13


def test_coplanar_and_far_segments():
    assert O.segment_pair_term([0, 0, 0], [1, 0, 0], [0, 2, 0], [1, 3, 0]) == 0.0
    v = O.segment_pair_term([0, 0, 0], [1e-3, 0, 0], [100, 0, 1], [100, 1e-3, 1])
    assert abs(v) < 1e-9


def test_invariance_rotation_translation():
    g = np.random.default_rng(7)
    p0, p1, q0, q1 = (g.normal(size=3) for _ in range(4))
    v = O.segment_pair_term(p0, p1, q0, q1)
    Q, _ = np.linalg.qr(g.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    t = g.normal(size=3) * 10
    w = O.segment_pair_term(*(Q @ x + t for x in (p0, p1, q0, q1)))
    assert abs(v - w) < 1e-12


def _hopf(n=200):
    t = 2 * np.pi * np.arange(n) / n
    a = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    b = np.stack([1 + np.cos(t), 0 * t, np.sin(t)], 1)
    return a, b


def _torus_link(n=400, m=2):
    t = 2 * np.pi * np.arange(n) / n
    out = []
    for k in range(2):
        th = m * t + k * np.pi
        out.append(np.stack([(2 + 0.5 * np.cos(th)) * np.cos(t), (2 + 0.5 * np.cos(th)) * np.sin(t),
                             0.5 * np.sin(th)], 1))
    return out


def test_whole_curve_sums_and_variants_agree():
    a, b = _hopf()
    v = O.ds(a, b)
    assert abs(abs(v) - 1) < 1e-9
    assert abs(O.ds(a[::-1], b) + v) < 1e-9          # orientation reversal
    assert abs(O.ds(b, a) - v) < 1e-9                # symmetry
    assert abs(O.ds_fast(a, b) - v) < 1e-9           # the two reference variants
    c, d = _torus_link()
    w = O.ds(c, d)
    assert abs(abs(w) - 2) < 1e-9 and abs(O.ds_fast(c, d) - w) < 1e-9
    far = b + [10.0, 0, 0]
    assert abs(O.ds(a, far)) < 1e-9


def test_bh_matches_ds_and_estimates_error():
    c, d = _torus_link(3000)
    exact = O.ds(c, d)
    for par in (True, False):
        v, err, vis = O.bh_pair(c, d, beta=2.0, parallel=par)
        assert round(v) == round(exact) and abs(v - exact) < 0.05
        assert vis > 0 and err >= 0
    # larger beta, smaller error
    v1, _, _ = O.bh_pair(c, d, beta=1.0)
    v4, _, _ = O.bh_pair(c, d, beta=4.0)
    assert abs(v4 - exact) < abs(v1 - exact)


def _brute_moments(p, segs, ctr):
    """Moments of the given segments about ctr by direct integration (2-point
    Gauss-Legendre is exact for these polynomials)."""
    n = len(p)
    c = np.zeros(3)
    D = np.zeros((3, 3))
    Q = np.zeros((3, 3, 3))
    gl = [0.5 - 0.5 / math.sqrt(3), 0.5 + 0.5 / math.sqrt(3)]
    for s in segs:
        a, b = p[s], p[(s + 1) % n]
        t = b - a
        c += t
        for u in gl:
            r = a + u * t - ctr
            D += 0.5 * np.outer(t, r)
            Q += 0.5 * t[:, None, None] * np.outer(r, r)[None]
    return c, D, Q


def test_tree_moments_brute_force():
    g = np.random.default_rng(3)
    p = np.cumsum(g.normal(size=(37, 3)), axis=0)
    mom, boxes, ch, seg = O.tree_moments(p)
    nb = len(boxes)

    def leaves(node):
        if node >= nb:
            return [int(seg[node - nb])]
        return leaves(ch[node][0]) + leaves(ch[node][1])

    for node in range(nb):
        ctr = 0.5 * (boxes[node][:3] + boxes[node][3:])
        c, D, Q = _brute_moments(p, leaves(node), ctr)
        m = mom[node]
        assert np.allclose(m[36:39], c, atol=1e-11)
        assert np.allclose(m[27:36].reshape(3, 3), D, atol=1e-10)
        assert np.allclose(m[0:27].reshape(3, 3, 3), Q, atol=1e-9)
        assert np.allclose(m[39:42], ctr)
    # leaf closed form: c = t, D = 0, Q_k = t_k t t^T / 12, |t|^3/36
    t = p[1] - p[0]
    k = nb + list(seg).index(0)
    assert np.allclose(mom[k][36:39], t)
    assert np.allclose(mom[k][0:27].reshape(3, 3, 3), t[:, None, None] * np.outer(t, t)[None] / 12)
    assert abs(mom[k][44] - np.linalg.norm(t) ** 3 / 36) < 1e-12


def _taylor_far_field(segs1, segs2, c1, c2):
    """Second-order Taylor expansion of the Gauss kernel about the centres,
    integrated over all segment pairs (independent of the moment algebra):
    (1/4 pi) sum int int g(R0 + d2 - d1) . (t2 x t1), g(R) = R/|R|^3.
    segs are lists of (start, end) points."""
    R0 = c2 - c1
    r = np.linalg.norm(R0)
    I = np.eye(3)
    J = I / r ** 3 - 3 * np.outer(R0, R0) / r ** 5
    H = (-3 * (np.einsum("ij,k->ijk", I, R0) + np.einsum("ik,j->ijk", I, R0) +
               np.einsum("jk,i->ijk", I, R0)) / r ** 5 +
         15 * np.einsum("i,j,k->ijk", R0, R0, R0) / r ** 7)
    g0 = R0 / r ** 3
    gl = [0.5 - 0.5 / math.sqrt(3), 0.5 + 0.5 / math.sqrt(3)]
    total = 0.0
    for a0, a1 in segs1:
        t1 = a1 - a0
        for b0, b1 in segs2:
            t2 = b1 - b0
            w = np.cross(t2, t1)
            for u in gl:
                for v in gl:
                    dl = (b0 + v * t2 - c2) - (a0 + u * t1 - c1)
                    gv = g0 + J @ dl + 0.5 * np.einsum("ijk,j,k->i", H, dl, dl)
                    total += 0.25 * gv @ w
    return total / (4 * math.pi)


def _node_segments(p, ch, seg, nb, node):
    if node >= nb:
        s = int(seg[node - nb])
        return [(p[s], p[(s + 1) % len(p)])]
    return _node_segments(p, ch, seg, nb, ch[node][0]) + _node_segments(p, ch, seg, nb, ch[node][1])


def test_far_field_equals_taylor_expansion_and_converges():
    """Nodes below the root hold open sub-chains, whose Gauss integral is not
    topologically zero, so the expansion and its error are both exercised."""
    g = np.random.default_rng(11)
    p1 = 0.3 * np.cumsum(g.normal(size=(24, 3)), axis=0)
    p2 = 0.3 * np.cumsum(g.normal(size=(30, 3)), axis=0)
    m1, b1, ch1, s1 = O.tree_moments(p1)
    m2, b2, ch2, s2 = O.tree_moments(p2)
    n1, n2 = len(b1), len(b2)
    A, B = ch1[n1 - 1][0], ch2[n2 - 1][1]
    assert A < n1 and B < n2
    segA, segB = _node_segments(p1, ch1, s1, n1, A), _node_segments(p2, ch2, s2, n2, B)
    rA = 0.5 * np.linalg.norm(b1[A][3:] - b1[A][:3])
    rB = 0.5 * np.linalg.norm(b2[B][3:] - b2[B][:3])
    prev = None
    for dist in (24.0, 48.0, 96.0):
        shift = np.array([dist, 0.3, -0.2])
        mB = m2[B].copy()
        mB[39:42] += shift
        ff, est = O.far_field(m1[A], mB, rA, rB)
        tay = _taylor_far_field(segA, [(a + shift, b + shift) for a, b in segB], m1[A][39:42], mB[39:42])
        assert abs(ff - tay) <= 1e-11 * abs(tay) + 1e-16
        exact = sum(O.segment_pair_term(a0, a1, b0 + shift, b1 + shift) for a0, a1 in segA for b0, b1 in segB)
        e = abs(exact - ff)
        assert e < 10 * est  # the Eq. 14 estimate has the size of the observed error
        if prev is not None:
            assert e < prev / 16  # the error falls like dist^-5 (asymptotic ratio 32)
        prev = e


def test_host_traversal_restatement_reproduces_oracle_bh():
    """tests/bh_host.py (which evaluates the device's exported trees on the
    GPU side) run over the oracle's own kd trees equals the oracle's
    Barnes–Hut: same visits, same value and estimate to FP64 summation
    order (frontiers stay below the reference's 75,000 BFS limit here)."""
    import bh_host as H
    from fastlinkingnumbers_amd import workloads as W
    for n, rg, beta in ((800, 6.0, 2.0), (600, 8.0, 1.5)):
        c, d = W.random_walk_pair(seed=4, n=n, rg=rg, workload=3)
        r = H.traverse(H.from_oracle(c), H.from_oracle(d), beta)
        v, est, vis = O.bh_pair(c, d, beta=beta)
        assert r["visits"] == vis
        assert abs(r["lk"] - v) < 1e-11 and abs(r["est"] - est) < 1e-11 * max(1.0, est)
        assert r["far"] > 0 and r["leaf_items"] > 0


def test_pls_tree_equals_brute_force_and_touching():
    g = np.random.default_rng(5)
    curves = []
    for k in range(120):
        c = g.uniform(0, 10, size=3)
        curves.append(c + g.uniform(-0.7, 0.7, size=(5, 3)))
    # two curves whose boxes touch exactly on a face
    curves.append(np.array([[20.0, 0, 0], [21.0, 1, 1], [20.5, 0.5, 0.2]]))
    curves.append(np.array([[21.0, 0, 0], [22.0, 1, 1], [21.5, 0.5, 0.2]]))
    m = O.Model.from_curves(curves)
    brute = m.pls(use_tree=False)
    assert np.array_equal(brute, m.pls(use_tree=True))
    assert [120, 121] in brute.tolist()


def test_bcc_roundtrip_and_obj_import(tmp_path):
    m = O.Model.load_bcc(os.path.join(DATA, "knittubeinit.bcc"))
    assert len(m) == 39 and sum(len(c) for c in m.curves()) == 18228
    assert m.info(0) == (O.BSPLINE, True)
    p = str(tmp_path / "rt.bcc")
    m.save_bcc(p)
    m2 = O.Model.load_bcc(p)
    assert all(np.array_equal(a, b) for a, b in zip(m.curves(), m2.curves()))
    for name in ("blenderobj", "houdiniobj"):
        mo = O.Model.load_obj(os.path.join(DATA, name + ".obj"))
        assert len(mo) == 2 and [len(c) for c in mo.curves()] == [4800, 4800]
    with pytest.raises(RuntimeError):
        O.Model.load_bcc(os.path.join(DATA, "blenderobj.obj"))


def test_discretize_invariants_polyline():
    """After discretization no segment box of a curve meets a segment box of
    a potentially linked curve, and every output vertex lies on its input
    polyline in order."""
    from fastlinkingnumbers_amd import workloads as W
    curves = W.chainmail(grid=3, nmin=24, nmax=40)
    m = O.Model.from_curves(curves)
    d = m.discretize()
    out = d.curves()
    pairs = m.pls()
    for i, j in pairs:
        bi = [np.r_[np.minimum(a, b), np.maximum(a, b)] for a, b in zip(out[i], np.roll(out[i], -1, 0))]
        bj = [np.r_[np.minimum(a, b), np.maximum(a, b)] for a, b in zip(out[j], np.roll(out[j], -1, 0))]
        for x in bi:
            for y in bj:
                assert not (np.all(x[:3] <= y[3:]) and np.all(y[:3] <= x[3:]))
    for c, o in zip(curves, out):
        assert len(o) >= len(c)
        assert np.array_equal(o[0], c[0])
    tri = m.certificate("directsum", discretize=False)[0]
    assert d.certificate("directsum", discretize=False)[0] == tri


def test_flk_library_exports_every_header_symbol():
    """libflk.so loads without a GPU and exports every function of include/*.h."""
    import re
    import ctypes
    lib_path = os.path.join(ROOT, "fastlinkingnumbers_amd", "lib", "libflk.so")
    if not os.path.exists(lib_path):
        subprocess.run(["make", "-C", os.path.join(ROOT, "fastlinkingnumbers_amd", "csrc")], check=True)
    lib = ctypes.CDLL(lib_path)
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            with open(os.path.join(inc, fn)) as f:
                txt = f.read()
            names |= set(re.findall(r"^\s*(?:[a-zA-Z_][\w\s\*]*?)\b(flk_\w+)\s*\(", txt, re.M))
    assert len(names) >= 16
    for n in sorted(names):
        assert hasattr(lib, n), n
    from fastlinkingnumbers_amd import FLK_EXPORTS
    assert {n for n, _, _ in FLK_EXPORTS} == names


def test_certificate_rounding_rules():
    """Host rounding of pair sums into certificate entries (reference
    linking_number_certificate.cpp:185-213): nearest integer with halves away
    from zero, zeros dropped, a non-finite sum kept (as INT_MIN), (min, max)
    curve order, sorted."""
    import fastlinkingnumbers_amd as flk
    pairs = [[4, 1], [0, 2], [0, 3], [5, 2], [1, 2], [3, 5], [2, 3], [0, 1]]
    lk = [0.49, 0.5, -0.5, 1.5, -2.5, 0.2, float("nan"), 2.9999]
    assert flk.certificate_from_lk(pairs, lk) == [
        (0, 1, 3), (0, 2, 1), (0, 3, -1), (1, 2, -3), (2, 3, -2 ** 31), (2, 5, 2)]
    assert flk.certificate_from_lk(np.zeros((0, 2), np.int32), np.zeros(0)) == []
