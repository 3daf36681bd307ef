"""Coordinate upload through the pinned staging path: FP64 input that FP32
holds exactly travels as FP32, any other FP64 input as FP64, FP32 input as
is.  Every route must leave the same FP64 vertices on the device -- read back
through a discretization without potential links, which returns each closed
polyline's vertices unchanged -- and give the same linking numbers."""
import numpy as np
import pytest

import oracle as O

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

NO_PAIRS = np.zeros((0, 2), np.int32)


def _walks(sizes, seed):
    g = np.random.default_rng(seed)
    return [W.random_walk(g, n, 50.0).astype(np.float32).astype(np.float64) for n in sizes]


@pytest.fixture
def restore_tune(engine):
    saved = engine.get_tune()
    yield
    engine.set_tune(**saved)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_upload_roundtrip_exact_and_inexact(engine, restore_tune, threads):
    """3.6M vertices span several 8 MB staging buffers per worker.  The second
    model has one value that FP32 cannot hold, near the end, so the narrowed
    upload is abandoned after most of its buffers went out and the model is
    sent again as FP64; both must come back bit for bit."""
    engine.set_tune(ingest_threads=threads)
    curves = _walks([1_500_000, 1_200_000, 900_000], seed=11)
    got, d = engine.discretize(curves, flk.SEG_POLYLINE, True, NO_PAIRS)
    assert d["rounds"] == 1 and d["overlaps"] == [0]
    assert all(np.array_equal(g, c) for g, c in zip(got, curves))
    odd = [c.copy() for c in curves]
    odd[2][-5, 1] += 2.0 ** -40 * (1.0 + abs(odd[2][-5, 1]))
    assert np.float32(odd[2][-5, 1]) != odd[2][-5, 1]
    got, _ = engine.discretize(odd, flk.SEG_POLYLINE, True, NO_PAIRS)
    assert all(np.array_equal(g, c) for g, c in zip(got, odd))
    # the discretized polylines are now the engine's curves
    assert engine.ncurves == 3 and list(engine.sizes) == [len(c) for c in odd]
    boxes = engine.curve_boxes()
    ref = np.array([np.concatenate([c.min(0), c.max(0)]) for c in odd])
    assert np.array_equal(boxes, ref)


def test_f32_and_f64_inputs_give_identical_results(engine):
    """The same FP32 values uploaded as float32 and as float64 give the same
    boxes and bit-identical linking numbers on both methods; an FP64 model
    perturbed below FP32 resolution takes the FP64 route and agrees with the
    FP64 oracle."""
    curves = W.hopf_torus_circles(seed=3)
    xyz, offs = flk.pack_curves(curves)
    xyz = xyz.astype(np.float32).astype(np.float64)
    pairs = np.array([[0, 1], [2, 3]], np.int32)
    out = {}
    for dt in (np.float64, np.float32):
        engine.set_curves_packed(xyz.astype(dt), offs)
        out[dt] = (engine.curve_boxes(),
                   engine.linking_numbers(pairs, method="directsum").lk,
                   engine.linking_numbers(pairs, method="barneshut", fallback_tol=0.0).lk)
    for a, b in zip(out[np.float64], out[np.float32]):
        assert np.array_equal(a, b)
    ref = np.array([O.ds(curves[0], curves[1]), O.ds(curves[2], curves[3])])
    np.testing.assert_allclose(out[np.float32][1], ref, atol=1e-5)
    g = np.random.default_rng(4)
    fine = xyz + g.uniform(-1, 1, xyz.shape) * 2.0 ** -36
    engine.set_curves_packed(fine, offs)
    parts = [fine.reshape(-1, 3)[offs[c]:offs[c + 1]] for c in range(len(curves))]
    assert np.array_equal(engine.curve_boxes(), np.array([np.concatenate([c.min(0), c.max(0)]) for c in parts]))
    r = engine.linking_numbers(pairs, method="directsum", fp64=1)
    np.testing.assert_allclose(r.lk, [O.ds(parts[0], parts[1]), O.ds(parts[2], parts[3])], atol=1e-9)


def test_tune_rejects_bad_thread_count(engine, restore_tune):
    with pytest.raises(flk.FlkError, match="ingest_threads"):
        engine.set_tune(ingest_threads=-1)


def test_f32_upload_edge_cases(engine):
    """Empty curves, a single vertex, 2-3 vertex curves and an all-empty model
    through the FP32 entry point give the same boxes and linking numbers as
    the same values through the FP64 one."""
    t = np.linspace(0, 2 * np.pi, 64, endpoint=False)
    circle = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    tiny = [np.zeros((0, 3)), np.array([[0.0, 0.0, 0.0]]), np.array([[0.0, 0, -1], [0.0, 0, 1]]),
            np.array([[0.5, 0, -1], [0.5, 0, 1], [2.0, 0, 0]])]
    xyz, offs = flk.pack_curves([circle] + tiny)
    xyz = xyz.astype(np.float32).astype(np.float64)
    pairs = np.array([[0, k] for k in range(1, 5)], np.int32)
    res = {}
    for dt in (np.float64, np.float32):
        engine.set_curves_packed(xyz.astype(dt), offs)
        res[dt] = (engine.curve_boxes(), engine.linking_numbers(pairs, method="directsum").lk)
    assert np.array_equal(res[np.float64][0], res[np.float32][0], equal_nan=True)
    assert np.array_equal(res[np.float64][1], res[np.float32][1])
    for dt in (np.float64, np.float32):
        engine.set_curves_packed(np.zeros((0, 3), dt), np.zeros(4, np.int64))
        assert engine.ncurves == 3 and engine.potential_links().shape == (0, 2)
