"""A Barnes–Hut certificate as one captured HIP graph, and the memset-free
device sort behind it.

With params.graph (the default) a Barnes–Hut or hybrid certificate -- the
direct-sum part below bh_direct_threshold, every tree build (radix sort,
levels, moments), both traversal passes with the adaptive-beta or
learned-criterion re-run decided on the device, the merge and the read-back
-- is one graph launch.  It must give what the eager launch sequence gives
(graph = 0): the same integers, the same node-pair visits, far fields, leaf
pairs and re-runs, and the same sums up to the order of the FP64 atomic
additions, on every replay, with other work in between.  The lattice
workload's integers are the FP32 direct sum's, which the oracle pins
elsewhere.
"""
import numpy as np
import pytest

flk = pytest.importorskip("fastlinkingnumbers_amd")
from fastlinkingnumbers_amd import criterion as crit  # noqa: E402
from fastlinkingnumbers_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

_COUNTS = ("bh_visits", "bh_far", "bh_leaf_pairs", "retries", "mlp_evals", "ds_seg_pairs")


def _same(a, b):
    assert np.all(np.round(a.lk) == np.round(b.lk))
    np.testing.assert_allclose(a.lk, b.lk, rtol=0, atol=1e-9)
    np.testing.assert_allclose(a.err, b.err, rtol=1e-9, atol=1e-15)
    assert np.array_equal(a.flags, b.flags)
    for k in _COUNTS:
        assert a.stats[k] == b.stats[k], k


@pytest.mark.parametrize("criterion", ["beta", "mlp"])
def test_lattice_hybrid_certificate_is_one_graph(engine, criterion):
    """Helical lattice at 1/50 size, the direct sum below 512^2 segment
    pairs and Barnes–Hut above, in one batch and in batches of 20k segments
    (several tree builds in the graph)."""
    engine.mlp_set_weights(*crit.load_weights())
    curves = W.workload(4, scale=0.02)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    ds = engine.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
    for ksegs in (0, 20):
        kw = dict(method="barneshut", criterion=criterion, bh_direct_threshold=512 * 512, fallback_tol=0.0,
                  bh_batch_ksegs=ksegs)
        eager = engine.linking_numbers(pairs, graph=0, **kw)
        assert eager.stats["graph_launches"] == 0 and eager.stats["ms_bh_build"] > 0
        runs = []
        for rep in range(3):
            runs.append(engine.linking_numbers(pairs, **kw))
            # other work between the replays: an eager direct sum of a few pairs
            engine.linking_numbers(pairs[:3], method="directsum", graph=0, fallback_tol=0.0)
        for g in runs:
            print(criterion, ksegs, "graph ms", g.stats["ms_kernel"], "eager ms", eager.stats["ms_total"],
                  "visits", g.stats["bh_visits"], "max |graph - ds|", np.max(np.abs(g.lk - ds.lk)))
            assert g.stats["graph_launches"] == 1
            assert g.stats["bh_host_passes"] == 0
            assert g.stats["ms_bh_build"] == 0.0 and g.stats["ms_kernel"] > 0
            _same(g, eager)
        direct = (runs[0].flags & flk.FLK_PAIR_DIRECT) != 0
        assert 10 < int(np.sum(~direct)) < len(pairs)
        assert np.all(np.round(runs[0].lk) == np.round(ds.lk))
        assert np.max(np.abs(runs[0].lk - ds.lk)) < 0.05


def test_graph_retry_and_overflow_fallback(engine):
    """Random walks at the default beta: the pair is re-run with a larger
    beta, decided by the device between the graph's two passes.  With level
    buffers too small (bh_level_capacity) the graph's pass overflows and the
    batch re-runs on the host-driven schedule with the same outcome."""
    curves = W.random_walk_pair(seed=1, n=20000, rg=30.0, workload=3)
    engine.set_curves(curves)
    kw = dict(method="barneshut", fallback_tol=0.0)
    eager = engine.linking_numbers([[0, 1]], graph=0, **kw)
    g = engine.linking_numbers([[0, 1]], **kw)
    assert g.stats["graph_launches"] == 1 and g.stats["retries"] == 1
    assert g.flags[0] & flk.FLK_PAIR_RETRY
    _same(g, eager)
    engine.set_tune(bh_level_capacity=4096)
    small = engine.linking_numbers([[0, 1]], **kw)
    assert small.stats["graph_launches"] == 1 and small.stats["bh_host_passes"] >= 1
    _same(small, eager)


def _scaled(curves):
    """Stored coordinates as the engine keeps them when the origin stays 0:
    the input times 2^-ceil(log2(2 d)), d the largest curve box diagonal."""
    d = max(float(np.linalg.norm(c.max(0) - c.min(0))) for c in curves)
    return [c * 2.0 ** -np.ceil(np.log2(2.0 * d)) for c in curves]


def test_tree_order_is_the_stable_sort_of_morton_keys(engine):
    """The segment order of a device tree is the stable radix sort of the
    segments' 30-bit Morton keys (10 bits per axis of the curve's box):
    recomputed here in FP64 with the kernel's operations, the order must be
    numpy's stable argsort exactly.  The curve retraces part of itself, so
    many keys are equal and stability is exercised."""
    rng = np.random.default_rng(11)
    walk = np.cumsum(rng.normal(size=(3000, 3)), axis=0)
    back = walk[::-1][1:-1]
    zig = np.repeat(walk[:200:10], 20, axis=0) + 1e-9 * rng.normal(size=(400, 3))
    c = np.concatenate([walk, back, zig])
    other = c[:50] + 1000.0
    engine.set_curves([c, other])
    nodes, order = engine.tree_moments(0, leaf_size=1)
    p = _scaled([c, other])[0]
    q = np.roll(p, -1, axis=0)
    m = 0.5 * (p + q)
    lo = p.min(0)
    hi = p.max(0)
    inv = np.where(hi > lo, 1024.0 / np.where(hi > lo, hi - lo, 1.0), 0.0)
    cell = np.clip((m - lo) * inv, 0.0, 1023.0).astype(np.uint64)

    def spread(v):
        v = v & np.uint64(0x3FF)
        v = (v | (v << np.uint64(16))) & np.uint64(0x0000FF0000FF)
        v = (v | (v << np.uint64(8))) & np.uint64(0x00F00F00F00F)
        v = (v | (v << np.uint64(4))) & np.uint64(0x0C30C30C30C3)
        v = (v | (v << np.uint64(2))) & np.uint64(0x249249249249)
        return v

    key = (spread(cell[:, 0]) << np.uint64(2)) | (spread(cell[:, 1]) << np.uint64(1)) | spread(cell[:, 2])
    expect = np.argsort(key, kind="stable")
    assert len(np.unique(key)) < len(key) - 500  # many ties
    assert np.array_equal(order, expect)


def test_graph_fp64_direct_part_and_judge(engine):
    """The hybrid graph with its direct-sum part in FP64, and the FP64 judge
    after a graph launch: a fallback tolerance below the Barnes–Hut error
    sends the Barnes–Hut pairs to the FP64 direct sum after the graph, whose
    results must equal a plain FP64 direct sum of the same pairs."""
    curves = W.workload(4, scale=0.01)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    kw = dict(method="barneshut", bh_direct_threshold=512 * 512)
    eager = engine.linking_numbers(pairs, graph=0, fp64=1, fallback_tol=0.0, **kw)
    g = engine.linking_numbers(pairs, fp64=1, fallback_tol=0.0, **kw)
    assert g.stats["graph_launches"] == 1
    _same(g, eager)
    ds64 = engine.linking_numbers(pairs, method="directsum", fp64=1, fallback_tol=0.0)
    direct = (g.flags & flk.FLK_PAIR_DIRECT) != 0
    assert direct.any() and (~direct).any()
    # the direct-sum part of the graph is the FP64 direct sum (its tasks split
    # the pairs by the batch's work, so the sums agree to rounding)
    np.testing.assert_allclose(g.lk[direct], ds64.lk[direct], rtol=0, atol=1e-12)
    judged = engine.linking_numbers(pairs, fallback_tol=1e-12, **kw)
    assert judged.stats["graph_launches"] == 1
    redone = (judged.flags & flk.FLK_PAIR_FALLBACK) != 0
    assert redone.sum() == judged.stats["fallbacks"] and redone[~direct].all()
    np.testing.assert_allclose(judged.lk[redone], ds64.lk[redone], rtol=0, atol=1e-9)


def test_graph_recaptured_after_scratch_reallocation(engine):
    """A certificate graph holds pointers into the Barnes–Hut scratch.  A call
    that grows that scratch (a tree export of a larger curve set) must make
    the next certificate capture a new graph rather than replay the old one;
    the results stay those of the eager launch."""
    small = W.random_walk_pair(seed=4, n=4000, rg=15.0, workload=3)
    engine.set_curves(small)
    kw = dict(method="barneshut", fallback_tol=0.0)
    eager = engine.linking_numbers([[0, 1]], graph=0, **kw)
    first = engine.linking_numbers([[0, 1]], **kw)
    _same(first, eager)
    big = W.random_walk_pair(seed=5, n=200000, rg=60.0, workload=3)
    engine.set_curves(small + big)
    engine.tree_moments(2, leaf_size=1)  # grows the tree buffers
    again = engine.linking_numbers([[0, 1]], **kw)
    assert again.stats["graph_launches"] == 1
    _same(again, eager)
    both = engine.linking_numbers([[0, 1], [2, 3]], **kw)
    eager_both = engine.linking_numbers([[0, 1], [2, 3]], graph=0, **kw)
    _same(both, eager_both)


def test_graph_replay_uses_the_weights_set_before_it(engine):
    """The learned level kernel copies the network from an image packed on
    the device after every upload of the weights, outside the captured
    graph.  A graph captured with one network and replayed after another is
    set must give what the eager launch gives with the new one: its visits,
    evaluations, re-runs and sums, not the old network's."""
    curves = W.random_walk_pair(seed=2, n=20000, rg=30.0, workload=3)
    engine.set_curves(curves)
    P, nm = crit.load_weights()
    kw = dict(method="barneshut", criterion="mlp", fallback_tol=0.0)
    engine.mlp_set_weights(P, nm)
    g1 = engine.linking_numbers([[0, 1]], **kw)
    _same(g1, engine.linking_numbers([[0, 1]], graph=0, **kw))
    # the output bias raised by 2: every predicted log10 error 2 larger, so
    # fewer node pairs pass as far fields
    P2 = P.copy()
    P2[-1] += 2.0
    engine.mlp_set_weights(P2, nm)
    g2 = engine.linking_numbers([[0, 1]], **kw)
    assert g2.stats["graph_launches"] == 1
    e2 = engine.linking_numbers([[0, 1]], graph=0, **kw)
    print("visits", g1.stats["bh_visits"], "->", g2.stats["bh_visits"], "retries", g2.stats["retries"])
    _same(g2, e2)
    assert g2.stats["bh_visits"] > g1.stats["bh_visits"]
    engine.mlp_set_weights(P, nm)
    _same(engine.linking_numbers([[0, 1]], **kw), g1)


def test_hybrid_branches_independent_of_the_direct_sum_grid(engine):
    """A hybrid certificate's direct-sum part runs as a graph branch beside
    the Barnes-Hut batches, on a share of its resident blocks.  The share only
    schedules: the same certificate at full residency (a tuned block count)
    and at one block per CU gives the same integers, visits and sums, and the
    eager sequence gives them too."""
    curves = W.workload(4, scale=0.02)
    engine.set_curves(curves)
    pairs = engine.potential_links()
    kw = dict(method="barneshut", bh_direct_threshold=512 * 512, fallback_tol=0.0)
    shared = engine.linking_numbers(pairs, **kw)
    direct = (shared.flags & flk.FLK_PAIR_DIRECT) != 0
    assert shared.stats["graph_launches"] == 1 and direct.any() and (~direct).any()
    _same(shared, engine.linking_numbers(pairs, graph=0, **kw))
    for bpc in (8, 1):
        engine.set_tune(ds_blocks_per_cu=bpc)
        g = engine.linking_numbers(pairs, **kw)
        assert g.stats["graph_launches"] == 1
        _same(g, shared)
