"""The FP64 direct-sum judge (fallback) and the NaN report, through the C ABI
and the drop-in CLI.

Every pair whose sum lands farther than fallback_tol from an integer, or is
not finite, is recomputed by the FP64 direct-sum kernel and flagged
FLK_PAIR_FALLBACK; a pair whose FP64 sum is NaN is also flagged FLK_PAIR_NAN,
and the certificate keeps it with the reference's message
(src/linking_number_certificate.cpp:185-213, :294-322).
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
FB, NAN = flk.FLK_PAIR_FALLBACK, flk.FLK_PAIR_NAN


@pytest.mark.parametrize("graph", [0, 1])
def test_judge_replaces_an_off_integer_barnes_hut_sum(engine, graph):
    """Barnes–Hut at beta = 1 (node spheres merely disjoint) lands far from
    the integer on this random-walk pair; the judge replaces it by the FP64
    direct sum, whose integer is the oracle's."""
    c, d = W.random_walk_pair(seed=2, n=5000, rg=15.0, workload=3)
    engine.set_curves([c, d])
    ref = O.ds(c, d)
    kw = dict(method="barneshut", beta_init=1.0, beta_limit=1.0, graph=graph)
    raw = engine.linking_numbers([[0, 1]], fallback_tol=0.0, **kw)
    off = abs(raw.lk[0] - round(raw.lk[0]))
    print(f"beta 1: {raw.lk[0]:.6f} (off-integer {off:.3f}), oracle direct sum {ref:.9f}")
    assert off > 0.05
    assert raw.stats["fallbacks"] == 0 and not raw.flags[0] & FB
    res = engine.linking_numbers([[0, 1]], fallback_tol=0.05, **kw)
    assert res.flags[0] & FB and not res.flags[0] & NAN and res.stats["fallbacks"] == 1
    assert res.stats["ms_fallback"] > 0.0
    assert abs(res.lk[0] - ref) < 1e-9 and round(res.lk[0]) == round(ref)
    f64 = engine.linking_numbers([[0, 1]], method="directsum", fp64=1, graph=graph).lk[0]
    assert abs(res.lk[0] - f64) < 1e-12


@pytest.mark.parametrize("graph", [0, 1])
@pytest.mark.parametrize("method", ["directsum", "barneshut"])
def test_judge_redoes_exactly_the_pairs_beyond_the_tolerance(engine, method, graph):
    """With a tolerance below the FP32 error the linked pairs are recomputed
    in FP64 and equal the oracle to 1e-9; the pairs within the tolerance
    keep their first result bit for bit."""
    curves = W.workload(1, scale=0.5)
    engine.set_curves(curves)
    pairs = np.array([[0, 1], [2, 3], [0, 2], [4, 5], [1, 3], [6, 7]], np.int32)
    ref = np.array([O.ds(curves[i], curves[j]) for i, j in pairs])
    raw = engine.linking_numbers(pairs, method=method, fallback_tol=0.0, graph=graph)
    tol = 1e-12
    redo = np.abs(raw.lk - np.round(raw.lk)) > tol
    assert redo[0] and redo[1]
    res = engine.linking_numbers(pairs, method=method, fallback_tol=tol, graph=graph)
    print(method, "redone", redo.astype(int), "flags", res.flags)
    assert np.array_equal((res.flags & FB) != 0, redo)
    assert res.stats["fallbacks"] == int(redo.sum())
    np.testing.assert_allclose(res.lk[redo], ref[redo], rtol=0, atol=1e-9)
    if method == "directsum":  # a fixed reduction order: bit for bit
        assert np.array_equal(res.lk[~redo], raw.lk[~redo])
    else:  # FP64 atomics: equal to summation order
        np.testing.assert_allclose(res.lk[~redo], raw.lk[~redo], rtol=0, atol=1e-12)
    assert not np.any(res.flags & NAN)


def test_nan_pair_is_flagged_and_kept(engine):
    """A NaN vertex makes the pair's sum NaN in FP32 and in the FP64 judge:
    flagged FLK_PAIR_NAN (after a fallback), other pairs unaffected, and the
    certificate keeps the pair as the reference does."""
    c = W.workload(1, scale=0.1)
    bad = c[1].copy()
    bad[3, 1] = np.nan
    engine.set_curves([c[0], bad, c[2], c[3]])
    pairs = np.array([[0, 1], [2, 3]], np.int32)
    res = engine.linking_numbers(pairs, method="directsum")
    assert np.isnan(res.lk[0]) and res.flags[0] & NAN and res.flags[0] & FB
    assert res.flags[1] == 0 and round(res.lk[1]) == -2
    assert flk.certificate_from_lk(pairs, res.lk) == [(0, 1, -2 ** 31), (2, 3, -2)]


def _cli(args):
    return subprocess.run([flk.VERIFYCURVES_BIN] + args, capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("method", ["directsum", "barneshut"])
def test_cli_reports_the_judge(tmp_path, method):
    """--fallback_tol below the method's error: every candidate pair is
    recomputed by the FP64 judge, the CLI says how many, and the certificate
    is the reference's."""
    out = tmp_path / "kb.txt"
    r = _cli([f"--input={os.path.join(GOLDEN, 'data', 'knittubebroken.bcc')}", f"--output={out}",
              f"--method={method}", "--fallback_tol=1e-12"])
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"FP64 direct-sum judge recomputed (\d+) pair\(s\)\.", r.stdout)
    assert m and int(m.group(1)) > 0, r.stdout
    assert out.read_text() == "39\n5,6,-2\n"
    r = _cli([f"--input={os.path.join(GOLDEN, 'data', 'knittubebroken.bcc')}", f"--output={out}",
              f"--method={method}"])
    assert "judge recomputed" not in r.stdout and out.read_text() == "39\n5,6,-2\n"


def test_cli_reports_a_nan_pair(tmp_path):
    """A polyline BCC whose curve holds a NaN coordinate: the reference's
    stderr message, and the pair kept in the certificate."""
    c = W.workload(1, scale=0.1)
    bad = c[1].copy()
    bad[3, 1] = np.nan
    p = str(tmp_path / "nan.bcc")
    O.Model.from_curves([c[0], bad, c[2], c[3]]).save_bcc(p)
    out = tmp_path / "nan.txt"
    r = _cli([f"--input={p}", f"--output={out}", "--method=directsum", "--nodiscretize"])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Linking number (0,1) is NaN." in r.stderr
    assert out.read_text() == "4\n0,1,-2147483648\n2,3,-2\n"
