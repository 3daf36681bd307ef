"""Host-side tools on the CPU: obj2bcc (reference src/obj2bcc.cpp, importer
src/model.cpp:124-264) turns the reference's OBJ fixtures into BCC files whose
curves equal the oracle's import of the same OBJ files, rounded to the BCC
format's FP32; the files carry the loop flags and the requested segment type.
Also the benchmark's one-GPU contract: asked for several GPUs it exits with a
message instead of a measurement."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN

flk = pytest.importorskip("fastlinkingnumbers_amd")
DATA = os.path.join(GOLDEN, "data")


@pytest.mark.parametrize("name", ["blenderobj", "houdiniobj"])
@pytest.mark.parametrize("seg", [("PL", O.POLYLINE), ("BS", O.BSPLINE)])
def test_obj2bcc_matches_oracle_import(tmp_path, name, seg):
    if not os.path.exists(flk.OBJ2BCC_BIN):
        pytest.skip("host tools not built")
    code, kind = seg
    out = str(tmp_path / f"{name}.bcc")
    r = subprocess.run([flk.OBJ2BCC_BIN, "--input", os.path.join(DATA, name + ".obj"), "--output", out,
                        "--segment_type", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = O.Model.load_bcc(out)
    ref = O.Model.load_obj(os.path.join(DATA, name + ".obj"), kind)
    assert len(got) == len(ref) == 2
    for c in range(len(ref)):
        assert got.info(c) == (kind, ref.info(c)[1])
        np.testing.assert_array_equal(got.curves()[c], ref.curves()[c].astype(np.float32).astype(np.float64))


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU; asked for more it exits with a message before
    touching a device."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300, cwd=root)
    assert r.returncode == 2, r.stdout + r.stderr
    assert "one GPU" in r.stdout + r.stderr
    assert r.stdout.strip() == "" or not r.stdout.lstrip().startswith("{")
