"""GPU parity of the direct-sum tile kernels (FP32 and FP64) against the oracle.

Tolerances: FP64 kernel vs the FP64 oracle within 1e-9 per pair; FP32 kernel
within 1e-5 per pair at these sizes (measured errors are below 3e-7, so a
biased angle polynomial or a wrong polarization term fails), and certificates
identical.  The reference's own answer for the fixtures is pinned through its
certificates.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def _all_pairs(n):
    return np.array([(i, j) for i in range(n) for j in range(i + 1, n)], dtype=np.int32)


def test_config1_all_pairs_fp32_fp64(engine):
    curves = W.workload(1)
    engine.set_curves(curves)
    pairs = _all_pairs(len(curves))
    ref = np.array([O.ds(curves[i], curves[j]) for i, j in pairs])
    r32 = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
    r64 = engine.linking_numbers(pairs, method="directsum", fp64=1)
    print("fp32 max err", np.abs(r32.lk - ref).max(), "fp64 max err", np.abs(r64.lk - ref).max())
    assert np.abs(r64.lk - ref).max() < 1e-9
    assert np.abs(r32.lk - ref).max() < 1e-5
    cert = flk.certificate_from_lk(pairs, r32.lk)
    assert cert == flk.certificate_from_lk(pairs, ref) == [(0, 1, -1), (2, 3, -2)]
    assert not np.any(r32.flags & flk.FLK_PAIR_FALLBACK)


def test_pair_order_and_swap_symmetry(engine):
    curves = W.workload(1, scale=0.1)
    engine.set_curves(curves)
    a = engine.linking_numbers([[0, 1], [1, 0], [2, 3], [3, 2]], method="directsum").lk
    assert abs(a[0] - a[1]) < 1e-4 and abs(a[2] - a[3]) < 1e-4
    assert round(a[0]) == -1 and round(a[2]) == -2


def test_pls_and_boxes_match_oracle(engine):
    curves = W.chainmail(grid=12, nmin=32, nmax=200)
    m = O.Model.from_curves(curves)
    engine.set_curves(curves)
    assert np.array_equal(engine.curve_boxes(), m.boxes())
    got = engine.potential_links()
    ref = m.pls()
    assert np.array_equal(got, ref)
    assert np.array_equal(m.pls(use_tree=True), ref)


def test_pls_touching_boxes_count():
    """Closed-interval overlap: boxes that only touch are potential links."""
    boxes = np.array([[0, 0, 0, 1, 1, 1], [1, 0, 0, 2, 1, 1], [2.5, 0, 0, 3, 1, 1],
                      [0, 1, 1, 0.5, 2, 2]], dtype=np.float64)
    eng = flk.Engine(0)
    got = eng.potential_links(boxes)
    eng.close()
    assert got.tolist() == [[0, 1], [0, 3]]


def test_uneven_batch_matches_oracle(engine):
    curves = W.chainmail(seed=3, grid=8, nmin=40, nmax=3000)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    res = engine.linking_numbers(pairs, method="directsum")
    ref = np.array([O.ds(curves[i], curves[j]) for i, j in pairs])
    print("pairs", len(pairs), "max err", np.abs(res.lk - ref).max())
    assert np.abs(res.lk - ref).max() < 1e-5
    assert flk.certificate_from_lk(pairs, res.lk) == flk.certificate_from_lk(pairs, ref)


def test_discretized_fixture_certificate(engine):
    """knittubebroken, discretized by the oracle: the GPU direct sum gives the
    reference certificate (README.md:86-95: '39' then '5,6,-2')."""
    m = O.Model.load_bcc(os.path.join(GOLDEN, "data", "knittubebroken.bcc")).discretize()
    curves = m.curves()
    engine.set_curves(curves)
    pairs = engine.potential_links()
    assert np.array_equal(pairs, m.pls())
    res = engine.linking_numbers(pairs, method="directsum")
    assert flk.certificate_from_lk(pairs, res.lk) == [(5, 6, -2)]
    assert np.max(np.abs(res.lk - np.round(res.lk))) < 0.05


def test_degenerate_curves(engine):
    """Empty and single-vertex curves contribute nothing and do not fault."""
    c = W.workload(1, scale=0.05)
    curves = [c[0], np.zeros((0, 3)), c[1], np.array([[5.0, 5.0, 5.0]])]
    engine.set_curves(curves)
    res = engine.linking_numbers([[0, 2], [0, 1], [1, 2], [3, 0]], method="directsum",
                                 fallback_tol=0.0)
    assert round(res.lk[0]) == -1
    assert res.lk[1] == 0.0 and res.lk[2] == 0.0
    assert np.isfinite(res.lk[3])


def test_tile_shapes_and_autotune(engine):
    """Every direct-sum and traversal launch shape gives the same certificate
    (sums agree to FP32 summation order); the autotuner times the candidates
    on the caller's pairs, keeps a valid shape and leaves the statistics of
    the last computation alone."""
    default = engine.get_tune()
    try:
        _check_shapes(engine)
    finally:
        engine.set_tune(**default)


def _check_shapes(engine):
    curves = W.workload(2, scale=0.05)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    base = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
    cert = flk.certificate_from_lk(pairs, base.lk)
    for R in (2, 4, 8):
        for bpc in (0, 2, 4):
            engine.set_tune(ds_rows_per_lane=R, ds_blocks_per_cu=bpc, ds_tasks_per_wave=4)
            r = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
            assert flk.certificate_from_lk(pairs, r.lk) == cert
            np.testing.assert_allclose(r.lk, base.lk, rtol=0, atol=2e-5)
    bh0 = engine.linking_numbers(pairs, method="barneshut", fallback_tol=0.0)
    for ipt in (2, 8):
        engine.set_tune(bh_items_per_thread=ipt)
        r = engine.linking_numbers(pairs, method="barneshut", fallback_tol=0.0)
        assert r.stats["bh_visits"] == bh0.stats["bh_visits"]
        np.testing.assert_allclose(r.lk, bh0.lk, rtol=0, atol=1e-9)
    with pytest.raises(RuntimeError):
        engine.set_tune(ds_rows_per_lane=5)
    before = engine.stats()
    t = engine.autotune(pairs, method="barneshut", bh_direct_threshold=300 * 300)
    print("tuned", t)
    assert t["ds_rows_per_lane"] in (2, 4, 8) and t["bh_items_per_thread"] in (2, 4, 8)
    assert t["ds_blocks_per_cu"] in (0, 4) and t["ds_tasks_per_wave"] in (8, 16)
    assert engine.get_tune() == t
    assert engine.stats() == before
    r = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
    assert flk.certificate_from_lk(pairs, r.lk) == cert


def test_graph_replays_match_eager(engine):
    """The batched direct sum is captured in a HIP graph and replayed for a
    repeated pair set.  Each replay, with other eager work and
    synchronisations in between, must reproduce the eagerly launched batch
    bit for bit; a new curve set must not replay a stale graph."""
    curves = W.workload(2, scale=0.05)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    eager = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0, graph=0).lk
    other = pairs[: max(1, len(pairs) // 3)]
    other_eager = engine.linking_numbers(other, method="directsum", fallback_tol=0.0, graph=0).lk
    for rep in range(4):
        r = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)  # capture, then replays
        assert np.array_equal(r.lk, eager), rep
        assert r.stats["ds_seg_pairs"] > 0
        # eager work of the program between replays
        o = engine.linking_numbers(other, method="directsum", fallback_tol=0.0, graph=0)
        assert np.array_equal(o.lk, other_eager)
        engine.linking_numbers(pairs[:2], method="barneshut", fallback_tol=0.0)
    # new curves (shifted copies): same pair indices, different sums expected
    moved = [c + np.float32(0.25) * (i % 2) for i, c in enumerate(curves)]
    engine.set_curves(moved)
    fresh = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0, graph=0).lk
    again = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0).lk
    assert np.array_equal(again, fresh)
    assert not np.array_equal(fresh, eager)


def test_tiny_linked_rings_beside_a_large_curve(engine):
    """The coordinate scale follows the largest curve (here a ring of radius
    500), so a Hopf pair of radius 1e-4 has relative vectors near 1e-9 of the
    unit: the products in the FP32 angle underflow and the kernel's rare path
    redoes them from factors scaled by 2^40.  The FP32 sum must still match
    the FP64 kernel and the oracle."""
    c = W.workload(1, scale=0.2)
    t = np.linspace(0, 2 * np.pi, 4000, endpoint=False)
    big = np.stack([2000 + 500 * np.cos(t), 500 * np.sin(t), 0 * t], 1)
    tiny = [1e-4 * c[0], 1e-4 * c[1]]
    engine.set_curves(tiny + [big])
    r32 = engine.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
    r64 = engine.linking_numbers([[0, 1]], method="directsum", fp64=1)
    ref = O.ds(tiny[0], tiny[1])
    print("fp32", r32.lk[0], "fp64", r64.lk[0], "oracle", ref)
    assert abs(r64.lk[0] - ref) < 1e-9 and abs(abs(ref) - 1) < 1e-6
    assert abs(r32.lk[0] - ref) < 1e-5
