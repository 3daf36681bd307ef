"""Device discretization (reference src/discretize.cpp:102-281) against the
oracle's restatement of the reference Discretize on the fixtures: the same
overlap count in every round, the same vertex count per curve and the same
vertices (both evaluate the reference's spline formulas in FP64)."""
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

flk = pytest.importorskip("fastlinkingnumbers_amd")

pytestmark = pytest.mark.gpu


def _model(name):
    path = os.path.join(GOLDEN, "data", name)
    return O.Model.load_obj(path) if name.endswith(".obj") else O.Model.load_bcc(path)


@pytest.mark.parametrize("name", ["glove.bcc", "knittubeinit.bcc", "knittubebroken.bcc", "blenderobj.obj",
                                  "houdiniobj.obj"])
def test_fixture_discretization_matches_oracle(engine, capfd, name):
    m = _model(name)
    curves = m.curves()
    info = [m.info(c) for c in range(len(m))]
    pairs = m.pls()  # potential links on the spline boxes (pinned against brute force on the CPU)
    capfd.readouterr()
    ref = m.discretize().curves()
    ref_overlaps = [int(x) for x in re.findall(r"Iteration \d+, (\d+) overlaps", capfd.readouterr().out)]
    got, d = engine.discretize(curves, [t for t, _ in info], [lp for _, lp in info], pairs)
    print(name, d["rounds"], d["overlaps"], d["segments"], d["pieces"])
    assert d["overlaps"] == ref_overlaps
    assert d["segments"] == sum(len(c) for c in curves)
    assert [len(c) for c in got] == [len(c) for c in ref]
    for a, b in zip(got, ref):
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * max(1.0, float(np.abs(b).max())))


def test_discretization_guards(engine):
    """A degenerate input segment fails the reference's guard (and names it);
    curves without potential links pass through as closed polylines; two
    circles passing 0.01 apart are bisected until their pieces separate."""
    t = np.linspace(0, 2 * np.pi, 40, endpoint=False)
    a = np.stack([np.cos(t), np.sin(t), 0 * t], 1)
    b = np.stack([1.99 + np.cos(t), 0 * t, np.sin(t)], 1)
    got, d = engine.discretize([a, b], flk.SEG_POLYLINE, True, np.zeros((0, 2), np.int32))
    assert d["rounds"] == 1 and d["overlaps"] == [0]
    assert all(np.array_equal(g, c) for g, c in zip(got, [a, b]))
    bad = a.copy()
    bad[7] = bad[6]
    with pytest.raises(flk.FlkError, match="Smallest input segment"):
        engine.discretize([bad, b], flk.SEG_POLYLINE, True, [[0, 1]])
    got, d = engine.discretize([a, b], flk.SEG_POLYLINE, True, [[0, 1]])
    assert d["rounds"] > 2 and d["overlaps"][0] > 0 and d["overlaps"][-1] == 0 and d["pieces"] > 80
    ref = O.Model.from_curves([a, b])
    assert [len(c) for c in got] == [len(c) for c in ref.discretize().curves()]
