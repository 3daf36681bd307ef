"""flk_linking_numbers_noalloc: the direct sum with every batch array in the
library's fixed per-device array (no device allocation).

It must give exactly what flk_linking_numbers gives when launched eagerly
(same plan, same kernels, fixed reduction order): on two calls in a row, the
second with a smaller batch (the array keeps the previous call's contents),
in FP32 and FP64, through the FP64 judge, from a second context on the same
GPU, and it must refuse batches beyond its limits, and other methods, with an
error.
"""
import numpy as np
import pytest

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.lk, b.lk)
    assert np.array_equal(a.flags, b.flags)
    assert a.stats["ds_seg_pairs"] == b.stats["ds_seg_pairs"]
    assert a.stats["fallbacks"] == b.stats["fallbacks"]


def test_noalloc_equals_the_allocating_entry_point(engine):
    curves = W.workload(2, scale=0.05)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    assert len(pairs) > 30
    small = pairs[: len(pairs) // 3]
    for kw in ({"fallback_tol": 0.0}, {"fp64": 1}, {"fallback_tol": 1e-12}):
        for p in (pairs, small):  # the second call is smaller than the first
            ref = engine.linking_numbers(p, method="directsum", graph=0, **kw)
            got = engine.linking_numbers_noalloc(p, **kw)
            _same(got, ref)
            if kw.get("fallback_tol") == 1e-12:
                assert got.stats["fallbacks"] > 0
    print("pairs", len(pairs), "max |lk|", float(np.abs(ref.lk).max()))
    # a second context on the same GPU shares the array (calls do not overlap here)
    other = flk.Engine(0)
    try:
        other.set_curves(curves)
        _same(other.linking_numbers_noalloc(small, fallback_tol=0.0),
              engine.linking_numbers(small, method="directsum", graph=0, fallback_tol=0.0))
    finally:
        other.close()
    # empty batch
    assert len(engine.linking_numbers_noalloc(np.zeros((0, 2), np.int32)).lk) == 0


def test_noalloc_limits_and_methods(engine):
    t = np.linspace(0, 2 * np.pi, 2100, endpoint=False)
    c = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    engine.set_curves([c, c + [1.0, 0.0, 0.0]])
    with pytest.raises(flk.FlkError, match="at most 65536 pairs"):
        engine.linking_numbers_noalloc(np.tile([[0, 1]], (65537, 1)))
    # 65,536 pairs of 2,100 x 2,100 segments: 17 row blocks each, past the task limit
    with pytest.raises(flk.FlkError, match="at most 1048576 tasks"):
        engine.linking_numbers_noalloc(np.tile([[0, 1]], (65536, 1)))
    with pytest.raises(flk.FlkError, match="direct sum only"):
        engine.linking_numbers_noalloc([[0, 1]], method="barneshut")
    # the context and the array stay usable after the refusals
    ref = engine.linking_numbers([[0, 1]], method="directsum", graph=0)
    _same(engine.linking_numbers_noalloc([[0, 1]]), ref)
