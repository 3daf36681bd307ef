"""The drop-in CLI (verifycurves, obj2bcc) against the reference's fixtures:
certificates byte-identical to results/reference_certificates, the README's
--comparewith example, and the discretization counts equal to the oracle's."""
import os
import re
import subprocess

import pytest

import oracle as O
from conftest import GOLDEN

flk = pytest.importorskip("fastlinkingnumbers_amd")
pytestmark = pytest.mark.gpu
DATA = os.path.join(GOLDEN, "data")
CERTS = os.path.join(GOLDEN, "reference_certificates")


def _run(args):
    r = subprocess.run([flk.VERIFYCURVES_BIN] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def _bcc(tmp_path, name):
    if name.endswith("obj"):
        out = str(tmp_path / f"{name}.bcc")
        r = subprocess.run([flk.OBJ2BCC_BIN, "--input", os.path.join(DATA, name + ".obj"),
                            "--output", out, "--segment_type", "PL"], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        return out
    return os.path.join(DATA, name + ".bcc")


def _disc_lines(text):
    return [l for l in text.splitlines() if re.match(r"(Iteration \d+|Original segment count)", l)]


@pytest.mark.parametrize("method", ["directsum", "barneshut"])
@pytest.mark.parametrize("name", ["glove", "knittubeinit", "knittubebroken", "blenderobj", "houdiniobj"])
def test_fixture_certificates(tmp_path, name, method):
    inp = _bcc(tmp_path, name)
    out = tmp_path / "cert.txt"
    stdout = _run([f"--input={inp}", f"--output={out}", f"--method={method}"])
    with open(os.path.join(CERTS, name + ".txt")) as f:
        assert out.read_text() == f.read()
    oracle_out = O.run_cli(inp, str(tmp_path / "oracle.txt"), method="directsum").stdout
    assert _disc_lines(stdout) == _disc_lines(oracle_out)


def test_readme_compare_example(tmp_path):
    dp, dc = tmp_path / "knittube.dp.txt", tmp_path / "knittube.dc.txt"
    _run([f"--input={os.path.join(DATA, 'knittubebroken.bcc')}", f"--output={tmp_path / 'kb.txt'}",
          f"--comparewith={os.path.join(CERTS, 'knittubeinit.txt')}", f"--diffpairs={dp}",
          f"--diffcurves={dc}"])
    assert dp.read_text() == "5,6,-2\n"
    assert dc.read_text() == "5\n6\n"


def test_invalid_method_and_extension(tmp_path):
    r = subprocess.run([flk.VERIFYCURVES_BIN, f"--input={os.path.join(DATA, 'glove.bcc')}",
                        f"--output={tmp_path / 'x.txt'}", "--method=fmm"], capture_output=True, text=True)
    assert r.returncode != 0 and "Invalid method" in r.stderr
    r = subprocess.run([flk.VERIFYCURVES_BIN, f"--input={os.path.join(DATA, 'blenderobj.obj')}",
                        f"--output={tmp_path / 'x.txt'}"], capture_output=True, text=True)
    assert r.returncode != 0


def test_polyline_no_discretize(tmp_path):
    """--discretize=false on a polyline BCC: certificate equals the oracle's."""
    from fastlinkingnumbers_amd import workloads as W
    curves = W.chainmail(grid=5, nmin=100, nmax=400)
    m = O.Model.from_curves(curves)
    p = str(tmp_path / "cm.bcc")
    m.save_bcc(p)
    ref = O.Model.load_bcc(p).certificate("directsum", discretize=False)[0]
    for method in ("directsum", "barneshut"):
        out = tmp_path / f"{method}.txt"
        _run([f"--input={p}", f"--output={out}", f"--method={method}", "--nodiscretize"])
        lines = out.read_text().splitlines()
        assert lines[0] == "25"
        assert [tuple(int(x) for x in l.split(",")) for l in lines[1:]] == ref


def test_learned_criterion_cli(tmp_path):
    out = tmp_path / "kb.txt"
    stdout = _run([f"--input={os.path.join(DATA, 'knittubebroken.bcc')}", f"--output={out}",
                   "--criterion=mlp"])
    assert out.read_text() == "39\n5,6,-2\n"
    assert "Mean Off-Integer Error" in stdout


def test_autotune_cli(tmp_path):
    out = tmp_path / "kb.txt"
    stdout = _run([f"--input={os.path.join(DATA, 'knittubebroken.bcc')}", f"--output={out}",
                   "--method=directsum", "--autotune"])
    assert out.read_text() == "39\n5,6,-2\n"
    assert "Autotuned shapes: direct sum" in stdout
