"""Coordinate upload through the pinned staging path: FP64 input that FP32
holds exactly travels as FP32, any other FP64 input as FP64, FP32 input as
is.  Every route must leave the same FP64 vertices on the device -- read back
through a discretization without potential links, which returns each closed
polyline's vertices unchanged -- and give the same linking numbers.  BCC files
stream straight into the curve set (flk_load_bcc) with the same result as the
oracle's import of the file."""
import os

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


def _boxes(curves):
    return np.array([np.r_[c.min(0), c.max(0)] if len(c) else np.r_[[np.inf] * 3, [-np.inf] * 3]
                     for c in curves])


@pytest.mark.parametrize("name", ["glove", "knittubeinit", "knittubebroken"])
def test_load_bcc_streams_fixture_and_discretizes_in_place(engine, name):
    """flk_load_bcc reads the reference's BCC fixtures into the curve set: the
    same segment type, loop flags and control points as the oracle's import of
    the file; discretizing that set in place gives the same polylines as the
    host-array route, and the certificate of the reference's fixture."""
    from conftest import GOLDEN
    path = os.path.join(GOLDEN, "data", name + ".bcc")
    m = O.Model.load_bcc(path)
    info = engine.load_bcc(path)
    n = len(m)
    assert info["segment_type"] == m.info(0)[0] and info["points"] == sum(len(c) for c in m.curves())
    assert list(info["looped"]) == [m.info(c)[1] for c in range(n)]
    assert np.array_equal(engine.curve_boxes(), _boxes(m.curves()))
    pairs = m.pls()  # potential links on the spline boxes
    types = [m.info(c)[0] for c in range(n)]
    d = engine.discretize_current(types, info["looped"], pairs)
    sizes = engine.sizes.copy()
    ref, dref = engine.discretize(m.curves(), types, info["looped"], pairs)
    assert d["overlaps"] == dref["overlaps"] and list(sizes) == [len(c) for c in ref]
    got = engine.potential_links()
    res = engine.linking_numbers(got, method="barneshut")
    with open(os.path.join(GOLDEN, "reference_certificates", name + ".txt")) as f:
        lines = f.read().split()
    want = [tuple(int(v) for v in l.split(",")) for l in lines[1:]]
    assert int(lines[0]) == n and flk.certificate_from_lk(got, res.lk) == want


def test_load_bcc_polyline_and_errors(engine, tmp_path):
    """A polyline BCC written by the oracle loads to the same curves and
    linking numbers as the host-array route; unreadable, foreign and truncated
    files fail with the reference's messages, a rejected header keeping the
    previous curve set."""
    curves = [c.astype(np.float32).astype(np.float64) for c in W.hopf_torus_circles(seed=7)]
    path = str(tmp_path / "hopf.bcc")
    O.Model.from_curves(curves).save_bcc(path)
    info = engine.load_bcc(path)
    assert info["segment_type"] == flk.SEG_POLYLINE and all(info["looped"])
    pairs = np.array([[0, 1], [2, 3]], np.int32)
    a = engine.linking_numbers(pairs, method="directsum").lk
    boxes = engine.curve_boxes()
    engine.set_curves(curves)
    assert np.array_equal(boxes, engine.curve_boxes())
    assert np.array_equal(a, engine.linking_numbers(pairs, method="directsum").lk)
    with pytest.raises(flk.FlkError, match="Error opening input file"):
        engine.load_bcc(str(tmp_path / "missing.bcc"))
    from conftest import GOLDEN
    with pytest.raises(flk.FlkError, match="Invalid header"):
        engine.load_bcc(os.path.join(GOLDEN, "data", "blenderobj.obj"))
    assert np.array_equal(boxes, engine.curve_boxes())  # previous set kept
    raw = open(path, "rb").read()
    cut = str(tmp_path / "cut.bcc")
    open(cut, "wb").write(raw[: len(raw) - 100])
    with pytest.raises(flk.FlkError, match="Invalid number of control points read"):
        engine.load_bcc(cut)
