"""Oracle parity on the generators of the two Barnes–Hut benchmark
workloads, at sizes the FP64 oracle finishes in seconds:

* the 10,000-loop helical lattice (here 200 loops, 2.1e5 vertices, 408
  candidate pairs) through the direct-sum/Barnes–Hut hybrid, with the beta
  criterion and with the learned criterion;
* the interlinked random-walk pair of the learned-criterion workload (radius
  of gyration 200; here 30,000 vertices each) through learned and beta
  Barnes–Hut.

Every pair is compared with the oracle's FP64 direct sum (the reference's
OneArctan/DirectFast sums, src/compute_link.cpp:126-280): identical integers,
|sum - oracle| < 0.05 for Barnes–Hut pairs, 1e-4 for direct-sum pairs (the
scaled-down lattice has 16-vertex rings whose long segments pass close to
their neighbours' segments, where FP32 terms lose a few more digits: measured
3.5e-5), and no pair needs the FP64 judge.
"""
import numpy as np
import pytest

import oracle as O

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import criterion as crit  # noqa: E402
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lattice():
    curves = W.workload(4, scale=0.02)
    xyz, off = flk.pack_curves(curves)
    pairs = O.Model.from_curves(curves).pls()
    return curves, pairs, O.ds_pairs(xyz, off, pairs)


@pytest.mark.parametrize("criterion", ["beta", "mlp"])
def test_helical_lattice_hybrid_matches_oracle(engine, lattice, criterion):
    curves, pairs, ref = lattice
    engine.mlp_set_weights(*crit.load_weights())
    engine.set_curves(curves)
    assert np.array_equal(engine.potential_links(), pairs)
    sizes = np.array([len(c) for c in curves], dtype=np.int64)
    thr = 512 * 512
    nbh = int(np.sum(sizes[pairs[:, 0]] * sizes[pairs[:, 1]] >= thr))
    res = engine.linking_numbers(pairs, method="barneshut", criterion=criterion, bh_direct_threshold=thr,
                                 fallback_tol=0.0)
    direct = (res.flags & flk.FLK_PAIR_DIRECT) != 0
    err = np.abs(res.lk - ref)
    print(f"{criterion}: {len(pairs)} pairs, {nbh} Barnes-Hut; max |bh - oracle| {err[~direct].max():.3g}, "
          f"max |ds - oracle| {err[direct].max():.3g}, linked {int(np.sum(np.abs(np.round(ref)) > 0))}, "
          f"visits {res.stats['bh_visits']}")
    assert int(np.sum(~direct)) == nbh > 100
    assert np.array_equal(np.round(res.lk), np.round(ref))
    worst = int(np.argmax(np.where(direct, err, -1.0)))
    print("worst direct-sum pair", pairs[worst], "sizes", sizes[pairs[worst]], "error", err[worst])
    assert err[~direct].max() < 0.05 and err[direct].max() < 1e-4
    assert flk.certificate_from_lk(pairs, res.lk) == flk.certificate_from_lk(pairs, ref)
    judged = engine.linking_numbers(pairs, method="barneshut", criterion=criterion, bh_direct_threshold=thr)
    assert judged.stats["fallbacks"] == 0


@pytest.mark.parametrize("criterion", ["mlp", "beta"])
def test_learned_workload_pair_matches_oracle(engine, criterion):
    curves = W.random_walk_pair(n=30000, rg=200.0, workload=5)
    ref = O.ds(*curves)
    assert abs(ref - round(ref)) < 1e-9 and round(ref) != 0
    engine.mlp_set_weights(*crit.load_weights())
    engine.set_curves(curves)
    res = engine.linking_numbers([[0, 1]], method="barneshut", criterion=criterion, fallback_tol=0.0)
    s = res.stats
    print(f"{criterion}: {res.lk[0]:.6f} oracle {ref:.9f} |err| {abs(res.lk[0] - ref):.3g} est {res.err[0]:.3g} "
          f"visits {s['bh_visits']} retries {s['retries']}")
    assert round(res.lk[0]) == round(ref)
    assert abs(res.lk[0] - ref) < 0.05


def test_near_approaches_in_a_tall_model(engine):
    """The five pairs of the full-size helical lattice that pass closest (their
    rings come within 3e-5 to 1e-4 of each other, one to three segment
    lengths, with up to 246,473 vertices) inside a model 160 units tall: the
    FP32 direct sum equals the FP64 direct sum.  With coordinates scaled by
    the model's height the FP32 sixth powers of those relative vectors
    underflowed and the sums landed up to 0.33 off their integers."""
    idx = [814, 915, 2155, 2156, 1691, 1692, 6385, 6486, 3647, 3648]
    rings = W.helical_lattice(only=idx)
    t = np.linspace(0, 2 * np.pi, 32, endpoint=False)
    cap = np.stack([np.cos(t), np.sin(t), 0 * t + 158.4], 1)  # the lattice's top, for its full height
    engine.set_curves(rings + [cap])
    pairs = np.arange(10, dtype=np.int32).reshape(5, 2)
    r32 = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
    r64 = engine.linking_numbers(pairs, method="directsum", fp64=1)
    print("fp32", r32.lk, "fp64", r64.lk)
    assert np.array_equal(np.round(r64.lk), [1, -1, -1, 0, -1])
    assert np.max(np.abs(r64.lk - np.round(r64.lk))) < 1e-9
    np.testing.assert_allclose(r32.lk, r64.lk, rtol=0, atol=1e-5)
