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


def _bench_module():
    import importlib.util
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(root, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_fan_pieces_sum_to_the_whole_curve():
    """The FP64 reference of the 2 x 8M pair splits curve 0 into closed
    sub-loops through vertex 0; their linking integrals with the other curve
    must add up to the whole curve's (the chords cancel), checked with the
    oracle's FP64 direct sum."""
    from fastlinkingnumbers_amd import workloads as W
    bench = _bench_module()
    a, b = W.random_walk_pair(seed=3, n=700, rg=6.0, workload=5)
    whole = O.ds(a, b)
    for k in (2, 5, 16):
        pieces = bench.fan_pieces(a, k)
        assert len(pieces) == k and sum(len(p) for p in pieces) == len(a) + 2 * (k - 1)
        assert abs(sum(O.ds(p, b) for p in pieces) - whole) < 1e-10


def test_bench_cpu_worker_runs_the_oracle():
    """The CPU baseline's child process: oracle only (no device), one JSON
    line with the run times and the pair sums, equal to the oracle's."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    spec = {"kind": "pairs", "workload": 2, "scale": 0.01, "reps": 2, "pairs": [[0, 1], [0, 2]]}
    import json
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--cpu-worker", json.dumps(spec)],
                       capture_output=True, text=True, timeout=300, cwd=root)
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert len(out["times"]) == 2 and len(out["lk"]) == 2
    from fastlinkingnumbers_amd import workloads as W
    curves = W.workload(2, scale=0.01)
    xyz, offs = flk.pack_curves(curves)
    np.testing.assert_allclose(out["lk"], O.ds_pairs(xyz, offs, np.array([[0, 1], [0, 2]], np.int32)), atol=1e-12)


def test_lattice_subset_equals_the_whole():
    """helical_lattice(only=...) builds the same rings as the whole lattice."""
    from fastlinkingnumbers_amd import workloads as W
    whole = W.helical_lattice(nloops=240, nmin=16, nmax=300)
    part = W.helical_lattice(nloops=240, nmin=16, nmax=300, only=[200, 3, 101])
    for k, ring in zip([200, 3, 101], part):
        assert np.array_equal(ring, whole[k])
