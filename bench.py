#!/usr/bin/env python3
"""Benchmark: certificate wall time and Gseg-pairs/s of the MI355X engine.

Default workload: the batched direct sum on the 40x40 chain-mail sheet (1,600
fp32 loops, vertex counts log-uniform in [256, 65536]).  One step = one pass of
the hot path over the resident curves: device potential-link search, the
direct sum of every candidate pair in one persistent launch, the FP64 judge for
any pair off-integer, and certificate rounding.  The curves are polylines
uploaded before timing (resident in HBM); there is no discretization step (the
reference's --discretize=false on polyline input).

Prints one JSON line.  `roofline` is for the direct-sum tile kernel
(58 algorithmic FLOP per segment pair, against the 157.3 TFLOP/s FP32 peak),
timed with HIP events inside the engine on its own stream; with --method
barneshut it is for the traversal level kernel against 8 TB/s, with the HBM
bytes that rocprofv3 counted for the same workload (profiles/traffic.json).
`cpu_baseline` times the oracle's restatement of the reference's OpenMP path
in a child process bound as the reference would run (OMP_PROC_BIND=spread,
OMP_PLACES=cores), median of 3: the direct-sum pair loop on a deterministic
sample of the same candidate pairs, or with --method barneshut the whole
certificate (potential links, Barnes-Hut with adaptive beta, rounding).

The Barnes-Hut legs (--method barneshut, and the learned-criterion comparison
on the 2 x 8M random walks that the default run appends) also run the full
FP32 direct sum on the same curves and report the largest difference.  The
learned-criterion comparison reports the median of five graph launches for
the beta and learned certificates, and (`cpu_reference_bh`) the reference's
own Barnes-Hut certificate of the same pair on the host, median of 3, beside
the GPU's whole certificates (potential links included) timed the same way.
--fp64-check instead runs the FP64 direct sum of the 2 x 8M pair (about 5.5
minutes), split into 16 closed sub-loops of the first curve whose added
chords cancel pairwise, and prints one JSON line with every estimate's error
against it.
"""
import argparse
import hashlib
import json
import os
import statistics
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_SEG_PAIR = 58          # canonical count (grid point, 3 dots, 2 denominators, det, x', y')
FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak
HBM_PEAK_GBPS = 8000.0          # MI355X published HBM3E bandwidth
WORKLOAD_NAMES = {
    1: "hopf_torus_circles (10 loops, 24k vertices, fp64)",
    2: "chainmail 40x40 (1600 loops, log-uniform 256..65536 vertices, fp32)",
    3: "random_walk_pair 2x2M (rg 100, fp32)",
    4: "helical_lattice pitch 1.6 (10000 loops, log-uniform 64..262144 vertices, fp32)",
    5: "random_walk_pair 2x8M (rg 200, fp32)",
}
# profiles/traffic.json entries measured for (workload, scale, criterion)
TRAFFIC_KEYS = {(2, 1.0, "directsum"): "ds_tile_f32/chainmail1600",
                (5, 1.0, "beta"): "bh_level/random_walks_2x8M_beta"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or len(os.sched_getaffinity(0))
    return model, threads


def cpu_worker(spec):
    """Child process of the CPU baseline: regenerates the workload, runs the
    oracle (never the engine: no GPU here) `reps` times and prints the times
    and results as one JSON line."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from fastlinkingnumbers_amd import workloads
    import fastlinkingnumbers_amd as flk
    curves = workloads.workload(spec["workload"], scale=spec["scale"])
    times, out = [], None
    if spec["kind"] == "certificate":
        m = O.Model.from_curves(curves)
        for _ in range(spec["reps"]):
            t = time.perf_counter()
            trip, _, _, cand, lk = m.certificate("barneshut", discretize=False)
            times.append(time.perf_counter() - t)
            log(f"cpu certificate run {len(times)}: {times[-1]:.2f} s")
        out = {"triplets": [list(map(int, x)) for x in trip], "lk": [float(v) for v in lk],
               "pairs": np.asarray(cand).tolist()}
    else:
        xyz, offs = flk.pack_curves(curves)
        sel = np.asarray(spec["pairs"], dtype=np.int32).reshape(-1, 2)
        for _ in range(spec["reps"]):
            t = time.perf_counter()
            lk = O.ds_pairs(xyz, offs, sel)
            times.append(time.perf_counter() - t)
            log(f"cpu direct-sum run {len(times)}: {times[-1]:.2f} s")
        out = {"lk": [float(v) for v in lk]}
    out["times"] = times
    print(json.dumps(out), flush=True)


def run_cpu_child(spec, timeout=600):
    env = dict(os.environ, OMP_PROC_BIND="spread", OMP_PLACES="cores")
    # the child's progress lines go straight to this process's stderr
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-worker", json.dumps(spec)],
                       stdout=subprocess.PIPE, text=True, env=env, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed with status {r.returncode} (its stderr is above)")
    return json.loads(r.stdout.strip().splitlines()[-1])


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--method", default="directsum", choices=["directsum", "barneshut"])
    ap.add_argument("--cpu-seconds", type=float, default=8.0,
                    help="target CPU time of one oracle baseline run (3 runs; 0 disables)")
    ap.add_argument("--learned", type=int, default=1,
                    help="also run learned BH vs GPU direct sum on workload 5 (2x8M)")
    ap.add_argument("--ds-sample", action="store_true",
                    help="workload-5 direct sum on a 1/16 row sample (x16) instead of in full")
    ap.add_argument("--ds-check", type=int, default=1,
                    help="--method barneshut: also run the full FP32 direct sum on the same pairs")
    ap.add_argument("--direct-threshold", type=int, default=None,
                    help="Barnes-Hut pairs with n1*n2 below this use the direct sum "
                         "(default 4096^2 for workload 4, else 0)")
    ap.add_argument("--criterion", default="beta", choices=["beta", "mlp"],
                    help="Barnes-Hut opening criterion for --method barneshut")
    ap.add_argument("--graph", type=int, default=1,
                    help="1: each certificate is captured once and replayed as a HIP graph; 0: launched eagerly "
                         "(per-kernel traces, the build / traversal split)")
    ap.add_argument("--fp64-check", action="store_true",
                    help="FP64 direct sum of the workload-5 pair against every estimate (one JSON line)")
    ap.add_argument("--cpu-worker", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_worker is not None:
        cpu_worker(json.loads(args.cpu_worker))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import fastlinkingnumbers_amd as flk
    from fastlinkingnumbers_amd import workloads

    if args.fp64_check:
        fp64_check(flk, workloads)
        return

    t0 = time.perf_counter()
    curves = workloads.workload(args.workload, scale=args.scale)
    xyz, offs = flk.pack_curves(curves)
    sizes = np.diff(offs)
    log(f"workload {args.workload}: {len(curves)} curves, {int(offs[-1])} vertices "
        f"(generated in {time.perf_counter() - t0:.1f}s)")

    eng = flk.Engine(0)
    log("device:", eng.device_name())
    # The workloads are FP32 data (held as float64 arrays of FP32 values).
    # Upload timings, each the second of two calls (the first allocates):
    # float32 arrays through flk_set_curves_f32, and the float64 arrays
    # through flk_set_curves, which narrows them to FP32 on the way.
    xyz32 = xyz.astype(np.float32)
    upload = {}
    for name, arr in (("f32", xyz32), ("f64", xyz)):
        for _ in range(2):
            t = time.perf_counter()
            eng.set_curves_packed(arr, offs)
            upload[name] = (time.perf_counter() - t) * 1e3
        log(f"upload ({name} input, pinned staging, FP32 over PCIe, device staging): "
            f"{upload[name]:.1f} ms for {arr.nbytes / 1e9:.2f} GB of input")
    del xyz32
    upload_ms = upload["f32"]
    if args.direct_threshold is None:
        args.direct_threshold = 4096 * 4096 if args.workload == 4 else 0
    params = flk.default_params(method=args.method, criterion=args.criterion,
                                bh_direct_threshold=args.direct_threshold, graph=args.graph)
    if args.criterion == "mlp":
        from fastlinkingnumbers_amd import criterion
        eng.mlp_set_weights(*criterion.load_weights())

    phase = {}

    def step():
        t = time.perf_counter()
        pairs = eng.potential_links()
        phase["pls"] = time.perf_counter() - t
        res = eng.linking_numbers(pairs, params)
        t = time.perf_counter()
        cert = flk.certificate_from_lk(pairs, res.lk)
        phase["round"] = time.perf_counter() - t
        return pairs, res, cert

    for _ in range(args.warmup):
        pairs, res, cert = step()
    times, kern = [], []
    for _ in range(args.steps):
        t = time.perf_counter()
        pairs, res, cert = step()
        times.append(time.perf_counter() - t)
        kern.append(res.stats["ms_ds_kernel"])
        log(f"step {len(times)}: {times[-1] * 1e3:.1f} ms (engine {res.stats['ms_total']:.1f} ms: direct-sum "
            f"kernel {kern[-1]:.1f} ms, bh build {res.stats['ms_bh_build']:.1f} ms, traverse "
            f"{res.stats['ms_bh_traverse']:.1f} ms, graph launches {res.stats['graph_launches']} "
            f"({res.stats['ms_kernel']:.1f} ms device), visits {res.stats['bh_visits']}, fallback "
            f"{res.stats['ms_fallback']:.1f} ms; pls {phase['pls'] * 1e3:.1f} ms, rounding "
            f"{phase['round'] * 1e3:.1f} ms)")
    st = res.stats
    seg_pairs = int(np.sum(sizes[pairs[:, 0]] * sizes[pairs[:, 1]])) if len(pairs) else 0
    ms_step = 1e3 * float(np.mean(times))
    off_int = np.abs(res.lk - np.round(res.lk))
    value = seg_pairs / (ms_step * 1e-3) * 1e-9

    ms_kernel = float(np.mean(kern))
    measured = profiled_traffic()
    traffic_key = TRAFFIC_KEYS.get((args.workload, args.scale,
                                    args.criterion if args.method == "barneshut" else "directsum"))
    tr = measured.get(traffic_key) if traffic_key else None
    traffic_src = ("rocprofv3 FETCH_SIZE and WRITE_SIZE runs of this workload, calibrated "
                   f"(profiles/traffic.json '{traffic_key}')") if tr else None
    ds_check = None
    if args.method == "barneshut":
        # the level kernels' device time: HIP events on every level launch of
        # one more certificate after the timed steps (per-launch events slow
        # a traversal by ~15%, so the timed steps run without them)
        tparams = flk.default_params(method=args.method, criterion=args.criterion, timing=1,
                                     bh_direct_threshold=args.direct_threshold)
        ms_level = eng.linking_numbers(pairs, tparams).stats["ms_bh_level"]
        per_visit = tr["bytes_per_certificate"] / tr["visits"] if tr else None
        gbps = st["bh_visits"] * per_visit / (ms_level * 1e-3) * 1e-9 if tr else None
        roofline = {
            "bound": "hbm", "kernel": "bh_level (far field inline), all launches of one certificate",
            "achieved": round(gbps, 1) if tr else None, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(gbps / HBM_PEAK_GBPS, 4) if tr else None,
            "traffic": tr["bytes_per_certificate"] if tr else None,
            "traffic_source": traffic_src,
            "bytes_per_visit_measured": round(per_visit, 1) if tr else None,
            "units": int(st["bh_visits"]), "ms_kernel": round(ms_level, 3),
            "kernel_timing": "HIP events on each level launch of a certificate run after the timed steps",
            "note": None if tr else "no counter profile of this workload: HBM bytes not measured, no fraction",
        }
        if args.ds_check and len(pairs):
            t = time.perf_counter()
            ds = eng.linking_numbers(pairs, method="directsum", fallback_tol=0.0)
            ds_check = {"ms": round((time.perf_counter() - t) * 1e3, 1),
                        "max_abs_diff": float(np.max(np.abs(ds.lk - res.lk))),
                        "same_certificate": flk.certificate_from_lk(pairs, ds.lk) == cert,
                        "ds_max_off_integer": float(np.max(np.abs(ds.lk - np.round(ds.lk))))}
            log(f"full FP32 direct sum on the same pairs: {ds_check}")
    else:
        achieved = FLOP_PER_SEG_PAIR * st["ds_seg_pairs"] / (ms_kernel * 1e-3) * 1e-12 if ms_kernel > 0 else None
        roofline = {
            "bound": "valu",
            "kernel": "ds_tile_f32",
            "achieved": round(achieved, 3) if achieved else None,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4) if achieved else None,
            "traffic": tr["bytes_per_launch"] if tr else None,
            "traffic_source": traffic_src,
            "flop_per_unit": FLOP_PER_SEG_PAIR,
            "units_per_launch": int(st["ds_seg_pairs"]),
            "ms_per_launch": round(ms_kernel, 3),
        }

    cpu = cpu_baseline(args, flk, pairs, sizes, res, cert) if args.cpu_seconds > 0 and len(pairs) else None
    learned = learned_vs_directsum(eng, flk, workloads, args) if args.learned else None
    out = {
        "metric": "certificate wall time + Gseg-pairs/s; learned BH vs plain GPU directsum, LKs exact",
        "value": round(value, 3),
        "unit": "Gseg-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "certificate_ms": round(ms_step, 3),
        "upload_ms": round(upload_ms, 1),
        "upload_ms_f64_input": round(upload["f64"], 1),
        "certificate_ms_with_upload": round(ms_step + upload_ms, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": WORKLOAD_NAMES.get(args.workload, str(args.workload)),
            "scale": args.scale,
            "method": args.method,
            "criterion": args.criterion if args.method == "barneshut" else None,
            "discretize": False,
            "curves": len(curves),
            "vertices": int(offs[-1]),
            "candidate_pairs": int(len(pairs)),
            "seg_pairs": seg_pairs,
            "linked_pairs": len(cert),
            "certificate_sha1": hashlib.sha1(repr(cert).encode()).hexdigest(),
            "max_off_integer": float(off_int.max()) if len(off_int) else 0.0,
            "fallbacks": int(st["fallbacks"]),
            "direct_threshold": args.direct_threshold,
            "graph": args.graph,
            "ds_seg_pairs": int(st["ds_seg_pairs"]),
            "bh_visits": int(st["bh_visits"]),
        },
        "roofline": roofline,
        "ds_check": ds_check,
        "cpu_baseline": cpu,
        "learned_bh": learned,
    }
    eng.close()
    print(json.dumps(out), flush=True)


def cpu_baseline(args, flk, pairs, sizes, res, cert):
    """The oracle's restatement of the reference's OpenMP path, in a child
    process (OMP_PROC_BIND=spread, OMP_PLACES=cores), median of 3 runs."""
    model, threads = cpu_info()
    base = {"cores": threads, "cpu_model": model, "kind": "port", "runs": 3,
            "binding": "OMP_PROC_BIND=spread OMP_PLACES=cores"}
    if args.method == "barneshut":
        if int(np.sum(sizes)) > 5_000_000:
            log("cpu baseline skipped: workload too large for the CPU Barnes-Hut in a bounded time")
            return None
        r = run_cpu_child({"kind": "certificate", "workload": args.workload, "scale": args.scale, "reps": 3})
        dt = statistics.median(r["times"])
        seg = int(np.sum(sizes[pairs[:, 0]] * sizes[pairs[:, 1]]))
        cpu_cert = sorted(tuple(x) for x in r["triplets"])
        return dict(base, value=round(seg / dt * 1e-9, 4), unit="Gseg-pairs/s (direct-equivalent)",
                    certificate_s=round(dt, 3), times_s=[round(x, 3) for x in r["times"]],
                    sample="whole certificate (potential links, Barnes-Hut with the reference's adaptive "
                           "beta, rounding) through the oracle's restatement of ComputeFromModel",
                    certificate_agrees=cpu_cert == sorted(cert))
    cost = sizes[pairs[:, 0]] * sizes[pairs[:, 1]]
    order = np.arange(len(pairs))
    # a small deterministic sample sizes the real one to the target time
    cal = order[:: max(1, len(pairs) // 64)][:8]
    r = run_cpu_child({"kind": "pairs", "workload": args.workload, "scale": args.scale, "reps": 1,
                       "pairs": pairs[cal].tolist()})
    rate = max(float(np.sum(cost[cal])) / max(r["times"][0], 1e-6), 1e6)
    stride = max(1, int(np.ceil(np.sum(cost) / (rate * args.cpu_seconds))))
    sel = order[::stride]
    r = run_cpu_child({"kind": "pairs", "workload": args.workload, "scale": args.scale, "reps": 3,
                       "pairs": pairs[sel].tolist()})
    dt = statistics.median(r["times"])
    v = np.array(r["lk"])
    rate = float(np.sum(cost[sel])) / dt
    return dict(base, value=round(rate * 1e-9, 4), unit="Gseg-pairs/s",
                times_s=[round(x, 3) for x in r["times"]],
                certificate_s_extrapolated=round(float(np.sum(cost)) / rate, 1),
                sample=f"every {stride}th candidate pair ({len(sel)} pairs, {int(np.sum(cost[sel]))} segment "
                       f"pairs) through the oracle's restatement of the reference direct-sum pair loop "
                       f"(DirectFast per pair, OpenMP over pairs); the whole certificate's time is "
                       f"extrapolated linearly in segment pairs",
                certificate_agrees=flk.certificate_from_lk(pairs[sel], v) ==
                flk.certificate_from_lk(pairs[sel], res.lk[sel]))


def profiled_traffic():
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (calibrated) for the profiled workloads, from profiles/traffic.json."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return {}
    with open(path) as fh:
        return json.load(fh)


def learned_vs_directsum(eng, flk, workloads, args):
    """Workload 5: two interlinked 8M-vertex random-walk loops.  Times the GPU
    FP32 direct sum (in full, or a 1/16 row sample scaled by 16 with
    --ds-sample: its cost is exactly proportional to rows), Barnes–Hut with
    the reference's beta criterion (adaptive beta as the reference), and
    Barnes–Hut with the learned criterion, and reports each estimate's
    difference from the full direct sum."""
    from fastlinkingnumbers_amd import criterion
    if not os.path.exists(criterion.DEFAULT_WEIGHTS):
        log("no learned-criterion weights; skipping the learned comparison")
        return None
    curves = workloads.workload(5)
    a, b = curves
    res = {"workload": "random_walk_pair 2x8M (rg 200, fp32)", "vertices": [len(a), len(b)]}
    if not args.ds_sample:
        eng.set_curves(curves)
        log("workload 5: full FP32 direct sum (about a minute)")
        r = eng.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
        res["ds_ms"] = round(r.stats["ms_total"], 1)
        res["ds_lk"] = float(r.lk[0])
        res["ds_timing"] = "full"
        log(f"workload 5 direct sum: {res['ds_ms']} ms, LK {res['ds_lk']:.9f}")
    else:
        eng.set_curves([a[::16], b])
        eng.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
        r = eng.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
        res["ds_ms"] = round(16 * r.stats["ms_total"] * len(a) / (16 * len(a[::16])), 1)
        res["ds_timing"] = "1/16 row sample (every 16th vertex of curve 1) x 16"
        eng.set_curves(curves)
    P, nm = criterion.load_weights()
    eng.mlp_set_weights(P, nm)
    for name, kw in (("beta", {}), ("learned", {"criterion": "mlp"})):
        eng.linking_numbers([[0, 1]], method="barneshut", **kw)  # warm-up (captures the graph)
        # five certificates, one graph launch each: the median is reported
        runs = [eng.linking_numbers([[0, 1]], method="barneshut", **kw) for _ in range(5)]
        ms_runs = [r.stats["ms_total"] for r in runs]
        r = runs[ms_runs.index(sorted(ms_runs)[2])]
        st = r.stats
        # the same certificate launched eagerly, for the build / traversal split
        eager = eng.linking_numbers([[0, 1]], method="barneshut", graph=0, **kw).stats
        dev = eng.linking_numbers([[0, 1]], method="barneshut", timing=1, **kw).stats  # kernel times
        res[name] = {"ms": round(st["ms_total"], 2), "ms_runs": [round(x, 2) for x in ms_runs], "lk": float(r.lk[0]),
                     "graph_launches": int(st["graph_launches"]), "ms_graph_device": round(st["ms_kernel"], 2),
                     "ms_eager": round(eager["ms_total"], 1),
                     "off_integer": float(abs(r.lk[0] - round(r.lk[0]))),
                     "err_estimate": float(r.err[0]),
                     "visits": int(st["bh_visits"]), "far": int(st["bh_far"]),
                     "leaf_pairs": int(st["bh_leaf_pairs"]), "mlp_evals": int(st.get("mlp_evals", 0)),
                     "retries": int(st["retries"]), "fallbacks": int(st["fallbacks"]),
                     "ms_build": round(eager["ms_bh_build"], 2), "ms_traverse": round(eager["ms_bh_traverse"], 2),
                     "ms_level_kernels": round(dev["ms_bh_level"], 2), "ms_mlp_kernels": round(dev["ms_mlp"], 2)}
        if "ds_lk" in res:
            res[name]["abs_diff_vs_ds"] = float(abs(r.lk[0] - res["ds_lk"]))
        log(f"workload 5 {name}: {res[name]}")
    res["learned_speedup_vs_ds"] = round(res["ds_ms"] / res["learned"]["ms"], 1)
    res["learned_speedup_vs_beta"] = round(res["beta"]["ms"] / res["learned"]["ms"], 2)
    if args.cpu_seconds > 0:
        res["cpu_reference_bh"] = cpu_bh_reference(eng, workloads)
    res["same_integer"] = round(res["beta"]["lk"]) == round(res["learned"]["lk"])
    if "ds_lk" in res:
        res["same_integer_as_ds"] = round(res["ds_lk"]) == round(res["learned"]["lk"]) == round(res["beta"]["lk"])
    return res


def cpu_bh_reference(eng, workloads):
    """The reference's own Barnes-Hut certificate of the 2 x 8M pair on the
    host (the oracle's restatement of ComputeFromModel: potential links,
    within-pair BFS then DFS under OpenMP, adaptive beta, rounding), median
    of 3 runs, beside the GPU beta and learned certificates of the same pair
    timed the same way (potential links, then the pair, wall clock)."""
    model, threads = cpu_info()
    gpu = {}
    for name, kw in (("beta", {}), ("learned", {"criterion": "mlp"})):
        eng.linking_numbers(eng.potential_links(), method="barneshut", **kw)
        ms = []
        for _ in range(3):
            t = time.perf_counter()
            pairs = eng.potential_links()
            r = eng.linking_numbers(pairs, method="barneshut", **kw)
            ms.append((time.perf_counter() - t) * 1e3)
        gpu[name] = {"ms": round(statistics.median(ms), 2), "lk": float(r.lk[0]), "pairs": len(pairs)}
    log("workload 5: CPU Barnes-Hut certificate, 3 runs")
    r = run_cpu_child({"kind": "certificate", "workload": 5, "scale": 1.0, "reps": 3}, timeout=900)
    dt = statistics.median(r["times"])
    cpu_lk = r["lk"][0] if r["lk"] else None
    out = {"kind": "port", "cores": threads, "cpu_model": model, "runs": 3,
           "binding": "OMP_PROC_BIND=spread OMP_PLACES=cores", "times_s": [round(x, 3) for x in r["times"]],
           "certificate_s": round(dt, 3), "lk": cpu_lk, "gpu_whole_certificate": gpu,
           "same_integer": cpu_lk is not None and
           round(cpu_lk) == round(gpu["beta"]["lk"]) == round(gpu["learned"]["lk"]),
           "speedup_beta": round(dt * 1e3 / gpu["beta"]["ms"], 1),
           "speedup_learned": round(dt * 1e3 / gpu["learned"]["ms"], 1),
           "sample": "whole certificate of the 2 x 8M random-walk pair through the oracle's restatement of "
                     "ComputeFromModel"}
    log(f"workload 5 CPU Barnes-Hut: {out}")
    return out


def fan_pieces(c, k):
    """Closed sub-loops of the closed polyline c whose sum is c: piece 0 is
    vertices [0, m] closed back to vertex 0, piece i > 0 is vertex 0 followed
    by vertices [i m, (i + 1) m] (the last ends at the final vertex) and closed
    back to vertex 0.  Each chord from vertex 0 is traversed once in each
    direction by consecutive pieces, so the chords' Gauss terms cancel and the
    pieces' linking integrals with any curve sum to c's."""
    n = len(c)
    cuts = [round(i * n / k) for i in range(k)] + [n - 1]
    out = [c[: cuts[1] + 1]]
    for i in range(1, k):
        out.append(np.concatenate([c[:1], c[cuts[i]: cuts[i + 1] + 1]]))
    return out


def fp64_check(flk, workloads):
    """FP64 direct sum of the 2 x 8M random-walk pair, as 16 closed sub-loops
    of curve 0 against curve 1 (one FP64 batch per sub-loop, ~21 s each, so
    progress is printed per piece), then the FP32 direct sum, Barnes-Hut with
    the beta criterion and with the learned criterion on the whole pair."""
    from fastlinkingnumbers_amd import criterion
    a, b = workloads.workload(5)
    pieces = fan_pieces(a, 16)
    eng = flk.Engine(0)
    log("device:", eng.device_name())
    eng.set_curves(pieces + [b])
    nb = len(pieces)
    total, ms = 0.0, 0.0
    for i in range(nb):
        r = eng.linking_numbers([[i, nb]], method="directsum", fp64=1)
        total += float(r.lk[0])
        ms += r.stats["ms_ds_kernel"]
        log(f"fp64 piece {i + 1}/{nb}: {r.lk[0]:+.12f} ({r.stats['ms_ds_kernel']:.0f} ms kernel)")
    eng.set_curves([a, b])
    out = {"workload": "random_walk_pair 2x8M (rg 200, fp32)", "vertices": [len(a), len(b)],
           "fp64_ds_lk": total, "fp64_ds_kernel_ms": round(ms, 1), "fp64_pieces": nb}
    r = eng.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
    out["fp32_ds"] = {"lk": float(r.lk[0]), "ms": round(r.stats["ms_total"], 1)}
    log(f"fp32 direct sum: {out['fp32_ds']}")
    eng.mlp_set_weights(*criterion.load_weights())
    for name, kw in (("beta", {}), ("learned", {"criterion": "mlp"})):
        r = eng.linking_numbers([[0, 1]], method="barneshut", fallback_tol=0.0, **kw)
        out[name] = {"lk": float(r.lk[0]), "ms": round(r.stats["ms_total"], 1), "err_estimate": float(r.err[0])}
    for name in ("fp32_ds", "beta", "learned"):
        out[name]["abs_err_vs_fp64"] = abs(out[name]["lk"] - total)
    out["integer"] = int(round(total))
    out["all_same_integer"] = len({int(round(out[k]["lk"])) for k in ("fp32_ds", "beta", "learned")}
                                  | {out["integer"]}) == 1
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
