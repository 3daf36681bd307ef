"""Edge cases through the C ABI on the GPU: empty and tiny inputs, ragged
batches, invalid arguments (errors, never faults), and pairs in either
order."""
import numpy as np
import pytest

import oracle as O

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def test_empty_model_and_no_pairs(engine):
    engine.set_curves([])
    assert engine.potential_links().shape == (0, 2)
    for method in ("directsum", "barneshut"):
        r = engine.linking_numbers(np.zeros((0, 2), np.int32), method=method)
        assert len(r.lk) == 0 and r.stats["seg_pairs"] == 0
    got, d = engine.discretize([], flk.SEG_POLYLINE, True, np.zeros((0, 2), np.int32))
    assert got == [] and d["segments"] == 0


def test_single_curve_and_tiny_curves(engine):
    """One curve has no candidate pairs; curves of 0-3 vertices linked against a
    circle give 0 (or the exact value) on both methods without faults."""
    t = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    circle = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    engine.set_curves([circle])
    assert engine.potential_links().shape == (0, 2)
    tiny = [np.zeros((0, 3)), np.array([[0.0, 0.0, 0.0]]), np.array([[0.0, 0, -1], [0.0, 0, 1]]),
            np.array([[0.5, 0, -1], [0.5, 0, 1], [2.0, 0, 0]])]
    engine.set_curves([circle] + tiny)
    pairs = np.array([[0, k] for k in range(1, 5)], np.int32)
    ref = np.array([O.ds(circle, c) if len(c) else 0.0 for c in tiny])
    for method, tol in (("directsum", 1e-5), ("barneshut", 0.02)):  # BH: expansion error
        r = engine.linking_numbers(pairs, method=method, fallback_tol=0.0)
        assert np.all(np.isfinite(r.lk))
        np.testing.assert_allclose(r.lk, ref, atol=tol)
    # the 2-vertex "curve" closes on itself (there and back): linking number 0;
    # the triangle threads the circle once
    assert abs(ref[2]) < 1e-9 and abs(abs(ref[3]) - 1) < 1e-9


def test_pair_order_and_duplicates(engine):
    """(i, j) and (j, i) give the same linking number (the direct sum to FP32
    rounding, Barnes-Hut to its expansion error: the traversal splits node 1
    on ties); a repeated pair is computed twice (bit-identical for the
    direct sum, to FP64 summation order for Barnes-Hut)."""
    c = W.workload(1, scale=0.1)
    engine.set_curves(c)
    for method, tol, rep in (("directsum", 1e-6, 0.0), ("barneshut", 0.02, 1e-9)):
        r = engine.linking_numbers([[0, 1], [1, 0], [0, 1], [2, 3], [3, 2]], method=method, fallback_tol=0.0)
        assert abs(r.lk[0] - r.lk[2]) <= rep
        assert abs(r.lk[0] - r.lk[1]) < tol and abs(r.lk[3] - r.lk[4]) < tol
        assert round(r.lk[0]) == round(r.lk[1]) == -1 and round(r.lk[3]) == round(r.lk[4]) == -2


def test_invalid_arguments_raise(engine):
    c = W.workload(1, scale=0.1)
    engine.set_curves(c)
    with pytest.raises(flk.FlkError, match="out of range"):
        engine.linking_numbers([[0, len(c)]])
    with pytest.raises(flk.FlkError, match="out of range"):
        engine.linking_numbers([[-1, 0]])
    with pytest.raises(flk.FlkError, match="invalid method"):
        engine.linking_numbers([[0, 1]], method=7)
    with pytest.raises(flk.FlkError, match="bh_batch_ksegs"):
        engine.linking_numbers([[0, 1]], method="barneshut", bh_batch_ksegs=-1)
    with pytest.raises(flk.FlkError):
        engine.set_curves_packed(np.zeros(6), np.array([0, 2, 1], np.int64))  # decreasing offsets
    # the context stays usable after errors
    engine.set_curves(c)
    assert round(engine.linking_numbers([[0, 1]], method="directsum").lk[0]) == -1


def test_ragged_batch_matches_oracle(engine):
    """A batch mixing curve lengths 3..5000 (every tile edge case of the 4-row
    and 256-row blocks) against the FP64 oracle, pair by pair."""
    rng = np.random.default_rng(11)
    sizes = [3, 5, 63, 64, 65, 255, 256, 257, 1023, 1024, 1025, 5000]
    curves = []
    for k, n in enumerate(sizes):
        t = np.sort(rng.uniform(0, 2 * np.pi, n))
        curves.append(np.stack([np.cos(t) + 0.9 * (k % 3), np.sin(t), 0.2 * np.sin(3 * t) + 0.1 * k], 1))
    engine.set_curves(curves)
    pairs = np.array([[i, j] for i in range(len(sizes)) for j in range(i + 1, len(sizes))], np.int32)
    ref = O.ds_pairs(*flk.pack_curves(curves), pairs)
    for R in (2, 4, 8):
        engine.set_tune(ds_rows_per_lane=R)
        r = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0, graph=0)
        np.testing.assert_allclose(r.lk, ref, rtol=0, atol=2e-5)
        assert flk.certificate_from_lk(pairs, r.lk) == flk.certificate_from_lk(pairs, ref)
