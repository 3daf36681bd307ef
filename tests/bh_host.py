"""Host restatement of the two-tree Barnes–Hut traversal over given trees.

Test infrastructure.  It follows the reference's breadth-first pass
(src/compute_link.cpp:695-841): per node pair, radii are half box diagonals;
two leaves are summed exactly with the leaf term (:576-606); a pair with
|r|^2 + eps > (r1 + r2)^2 beta^2 + eps takes the second-order far field
(:410-524, Eq. 14 estimate); otherwise the node with the larger radius is
split (ties split node 1, a leaf is never split).  The far field and the leaf
term are the oracle's FP64 restatements (pinned against an independent Taylor
integration and the segment-pair quadrature), so a traversal of the device's
own exported trees here is an independent evaluation of what the device
traversal must produce.

A tree is a dict of per-node arrays: center (N,3), radius (N,), rsq (N,),
child (N,2) with -1 for a leaf, packed (N,45) moments in the oracle's layout
(Q[a][i][j] 27, D[i][j] 9, c 3, centre 3, cn, dn, qn), and per leaf node the
midpoints and tangents of its segments; node 0 is the root.
"""
import numpy as np

import oracle as O


def _pack(Q, D, c, center, cn, dn, qn):
    return np.concatenate([np.asarray(Q).reshape(27), np.asarray(D).reshape(9), c, center, [cn, dn, qn]])


def _seg_mid_tan(pts, segs):
    n = len(pts)
    a = pts[segs]
    b = pts[(segs + 1) % n]
    return 0.5 * (a + b), b - a


def from_device(nodes: dict, order: np.ndarray, pts: np.ndarray) -> dict:
    """Tree from Engine.tree_moments (input coordinates) of the closed
    polyline pts.  The device keeps each centre as an FP32 hi/lo pair; hi and
    lo are recovered for the FP32 classification (exact when the engine's
    origin is 0: its coordinate scale is a power of two)."""
    pts = np.asarray(pts, dtype=np.float64)
    N = len(nodes["radius"])
    packed = np.zeros((N, 45))
    leaves = {}
    for k in range(N):
        packed[k] = _pack(nodes["Q"][k], nodes["D"][k], nodes["c"][k], nodes["center"][k], nodes["cn"][k],
                          nodes["dn"][k], nodes["qn"][k])
        if nodes["child"][k][0] < 0:
            s0, ns = int(nodes["seg0"][k]), int(nodes["nseg"][k])
            leaves[k] = _seg_mid_tan(pts, order[s0:s0 + ns].astype(np.int64))
    hi = nodes["center"].astype(np.float32)
    lo = (nodes["center"] - hi.astype(np.float64)).astype(np.float32)
    return {"center": nodes["center"].copy(), "radius": nodes["radius"].copy(), "rsq": nodes["rsq"].copy(),
            "child": nodes["child"].copy(), "packed": packed, "leaves": leaves, "hi": hi, "lo": lo}


def _fma32(a, b, c):
    # float32 fused multiply-add: the float64 product of two float32 values
    # is exact, so one rounding of the sum (then to float32) stands in for it
    return np.float32(np.float64(a) * np.float64(b) + np.float64(c))


def _far_fp32(A, a, B, b, b2):
    """The device's FP32 far test (bh_level): r from the hi/lo centre
    differences, |r|^2 by fused multiply-adds, (r1 + r2)^2 beta^2 in FP32."""
    r = (B["hi"][b] - A["hi"][a]) + (B["lo"][b] - A["lo"][a])  # float32 arithmetic
    dsq = _fma32(r[0], r[0], _fma32(r[1], r[1], np.float32(r[2] * r[2])))
    rs = np.float32(np.float32(A["radius"][a]) + np.float32(B["radius"][b]))
    return bool(dsq > np.float32(np.float32(rs * rs) * np.float32(b2)))


def from_oracle(pts: np.ndarray) -> dict:
    """The oracle's kd tree of the closed polyline pts (the reference's
    MakeSegmentTrees + BuildTreeMoments), renumbered so the root is node 0."""
    pts = np.asarray(pts, dtype=np.float64)
    mom, boxes, ch, seg = O.tree_moments(pts)
    nb = len(boxes)
    N = 2 * nb + 1
    # oracle ids: internal 0..nb-1 (root nb-1), leaves nb..2nb
    ids = [nb - 1] + [k for k in range(N) if k != nb - 1]
    new = {old: i for i, old in enumerate(ids)}
    center = np.zeros((N, 3))
    rsq = np.zeros(N)
    child = -np.ones((N, 2), dtype=np.int64)
    packed = np.zeros((N, 45))
    leaves = {}
    for old, i in new.items():
        packed[i] = mom[old]
        center[i] = mom[old][39:42]
        if old < nb:
            lo, hi = boxes[old][:3], boxes[old][3:]
            child[i] = [new[int(ch[old][0])], new[int(ch[old][1])]]
        else:
            s = int(seg[old - nb])
            a, b = pts[s], pts[(s + 1) % len(pts)]
            lo, hi = np.minimum(a, b), np.maximum(a, b)
            leaves[i] = (mom[old][39:42][None], mom[old][36:39][None])
        rsq[i] = 0.25 * np.sum((hi - lo) ** 2)
    return {"center": center, "radius": np.sqrt(rsq), "rsq": rsq, "child": child, "packed": packed,
            "leaves": leaves}


def traverse(T1: dict, T2: dict, beta: float, eps_sq: float = 1e-24, fp32_test: bool = False) -> dict:
    """Breadth-first traversal from the root pair.  Returns the summed
    linking integral, the summed Eq. 14 estimate (unscaled), the visit, far
    field and leaf-item counts, the leaf segment pairs and the sum of the
    absolute values of the far-field terms (a scale for FP32 tolerances).
    fp32_test: decide far/split with the device's FP32 test (device trees),
    so the same node pairs are visited; the terms stay FP64."""
    b2 = beta * beta
    level = [(0, 0)]
    total = est = absum = 0.0
    visits = far = leaf_items = leaf_pairs = 0
    while level:
        nxt = []
        for a, b in level:
            visits += 1
            r = T2["center"][b] - T1["center"][a]
            dsq = float(r @ r) + eps_sq
            la, lb = T1["child"][a][0] < 0, T2["child"][b][0] < 0
            rs = T1["radius"][a] + T2["radius"][b]
            if la and lb:
                leaf_items += 1
                m1, t1 = T1["leaves"][a]
                m2, t2 = T2["leaves"][b]
                for i in range(len(m1)):
                    for j in range(len(m2)):
                        total += O.bh_leaf_term(m1[i], t1[i], m2[j], t2[j])
                leaf_pairs += len(m1) * len(m2)
            elif (_far_fp32(T1, a, T2, b, b2) if fp32_test else dsq > rs * rs * b2 + eps_sq):
                far += 1
                v, e = O.far_field(T1["packed"][a], T2["packed"][b], T1["radius"][a], T2["radius"][b], eps_sq)
                total += v
                est += e
                absum += abs(v)
            else:
                split_a = (not (T2["rsq"][b] > T1["rsq"][a])) if (not la and not lb) else lb
                if split_a:
                    nxt += [(int(T1["child"][a][0]), b), (int(T1["child"][a][1]), b)]
                else:
                    nxt += [(a, int(T2["child"][b][0])), (a, int(T2["child"][b][1]))]
        level = nxt
    return {"lk": total, "est": est, "visits": visits, "far": far, "leaf_items": leaf_items,
            "leaf_pairs": leaf_pairs, "far_abs": absum}
