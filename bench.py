#!/usr/bin/env python3
"""Benchmark: certificate wall time and Gseg-pairs/s of the MI355X engine.

Default workload: the batched direct sum on the 40x40 chain-mail sheet (1,600
fp32 loops, vertex counts log-uniform in [256, 65536]).  One step = one pass of
the hot path over the resident curves: device potential-link search, the
direct sum of every candidate pair in one persistent launch, the FP64 judge for
any pair off-integer, and certificate rounding.  The curves are uploaded before
timing (resident in HBM).

Prints one JSON line.  `roofline` is for the direct-sum tile kernel
(58 algorithmic FLOP per segment pair, against the 157.3 TFLOP/s FP32 peak),
timed with HIP events inside the engine on its own stream; with --method
barneshut it is for the traversal level kernel (336 canonical bytes per
node-pair visit against 8 TB/s).  `traffic` is the HBM bytes per launch that
rocprofv3 measured for the same workload (profiles/traffic.json).
`cpu_baseline` times the oracle's restatement of the reference's OpenMP
direct-sum pair loop on a deterministic sample of the same candidate pairs on
this host (with --method barneshut: the reference Barnes-Hut, whole workload).
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOP_PER_SEG_PAIR = 58          # canonical count (grid point, 3 dots, 2 denominators, det, x', y')
FP32_PEAK_TFLOPS = 157.3        # MI355X published FP32 vector peak
WORKLOAD_NAMES = {
    1: "hopf_torus_circles (10 loops, 24k vertices, fp64)",
    2: "chainmail 40x40 (1600 loops, log-uniform 256..65536 vertices, fp32)",
    3: "random_walk_pair 2x2M (rg 100, fp32)",
    4: "helical_lattice (10000 loops, log-uniform 64..262144 vertices, fp32)",
    5: "random_walk_pair 2x8M (rg 200, fp32)",
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", type=int, default=2)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--method", default="directsum", choices=["directsum", "barneshut"])
    ap.add_argument("--cpu-seconds", type=float, default=12.0,
                    help="target CPU time of the oracle baseline sample (0 disables)")
    ap.add_argument("--learned", type=int, default=1,
                    help="also run learned BH vs GPU direct sum on workload 5 (2x8M)")
    ap.add_argument("--full-ds", action="store_true",
                    help="time the full workload-5 direct sum instead of a 1/16 row sample")
    ap.add_argument("--direct-threshold", type=int, default=None,
                    help="Barnes-Hut pairs with n1*n2 below this use the direct sum "
                         "(default 4096^2 for workload 4, else 0)")
    ap.add_argument("--criterion", default="beta", choices=["beta", "mlp"],
                    help="Barnes-Hut opening criterion for --method barneshut")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}).",
              file=sys.stderr)
        sys.exit(2)

    import fastlinkingnumbers_amd as flk
    from fastlinkingnumbers_amd import workloads

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
                                bh_direct_threshold=args.direct_threshold)
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
            f"{res.stats['ms_bh_traverse']:.1f} ms, visits {res.stats['bh_visits']}, fallback "
            f"{res.stats['ms_fallback']:.1f} ms; pls {phase['pls'] * 1e3:.1f} ms, rounding "
            f"{phase['round'] * 1e3:.1f} ms)")
    st = res.stats
    seg_pairs = int(np.sum(sizes[pairs[:, 0]] * sizes[pairs[:, 1]])) if len(pairs) else 0
    ms_step = 1e3 * float(np.mean(times))
    off_int = np.abs(res.lk - np.round(res.lk))
    value = seg_pairs / (ms_step * 1e-3) * 1e-9

    ms_kernel = float(np.mean(kern))
    achieved = FLOP_PER_SEG_PAIR * st["ds_seg_pairs"] / (ms_kernel * 1e-3) * 1e-12 if ms_kernel > 0 else None
    measured = profiled_traffic()
    if args.method == "barneshut":
        # canonical bytes per visit over the level kernels' device time: HIP
        # events on every level launch of one more certificate after the timed
        # steps (per-launch events slow a traversal by ~15%, so the timed
        # steps run without them)
        tparams = flk.default_params(method=args.method, criterion=args.criterion, timing=1,
                                     bh_direct_threshold=args.direct_threshold)
        ms_level = eng.linking_numbers(pairs, tparams).stats["ms_bh_level"]
        gbps = st["bh_visits"] * BH_BYTES_PER_VISIT / (ms_level * 1e-3) * 1e-9
        t = measured.get("bh_level/random_walks_2x8M_beta") if (args.workload, args.scale, args.criterion) == (5, 1.0, "beta") else None
        roofline = {
            "bound": "hbm", "kernel": "bh_level (far field inline), all launches of one certificate",
            "achieved": round(gbps, 1), "peak": 8000.0, "unit": "GB/s",
            "frac": round(gbps / 8000.0, 4),
            "traffic": t["bytes_per_certificate"] if t else None,
            "traffic_GBps": round(t["bytes_per_certificate"] / (ms_level * 1e-3) * 1e-9, 1) if t else None,
            "bytes_per_unit": BH_BYTES_PER_VISIT, "units": int(st["bh_visits"]),
            "ms_kernel": round(ms_level, 3), "launches": "one per level chunk",
            "kernel_timing": "HIP events on each level launch of a certificate run after the timed steps",
        }
    else:
        t = measured.get("ds_tile_f32/chainmail1600") if (args.workload, args.scale) == (2, 1.0) else None
        roofline = {
            "bound": "valu",
            "kernel": "ds_tile_f32",
            "achieved": round(achieved, 3) if achieved else None,
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP32_PEAK_TFLOPS, 4) if achieved else None,
            "traffic": t["bytes_per_launch"] if t else None,
            "flop_per_unit": FLOP_PER_SEG_PAIR,
            "units_per_launch": int(st["ds_seg_pairs"]),
            "ms_per_launch": round(ms_kernel, 3),
        }

    cpu = None
    if args.cpu_seconds > 0 and len(pairs) and args.method == "barneshut":
        # the reference's Barnes-Hut (within-pair BFS+DFS for < 6 curves) on
        # the whole workload when it is small enough to finish in seconds
        if int(offs[-1]) <= 5_000_000:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            import oracle as O
            t = time.perf_counter()
            v = [O.bh_pair(curves[i], curves[j], beta=2.0, parallel=len(curves) < 6)[0] for i, j in pairs]
            dt = time.perf_counter() - t
            cpu = {"value": round(seg_pairs / dt * 1e-9, 4), "unit": "Gseg-pairs/s (direct-equivalent)",
                   "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), "kind": "port",
                   "sample": f"whole workload through the oracle's restatement of the reference Barnes-Hut "
                             f"(beta 2, no adaptive retry), {dt:.1f} s",
                   "certificate_agrees": flk.certificate_from_lk(pairs, np.array(v)) == cert}
        else:
            log("cpu baseline skipped: workload too large for the CPU Barnes-Hut in a bounded time")
    elif args.cpu_seconds > 0 and len(pairs):
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # CPU oracle: the baseline leg only
        cost = sizes[pairs[:, 0]] * sizes[pairs[:, 1]]
        order = np.arange(len(pairs))
        # calibration on a small deterministic sample, then a sample sized to the target time
        def run(sel):
            t = time.perf_counter()
            v = O.ds_pairs(xyz, offs, pairs[sel])
            return v, time.perf_counter() - t
        stride = max(1, len(pairs) // 64)
        sel = order[::stride][:8]
        _, dt = run(sel)
        rate = max(float(np.sum(cost[sel])) / max(dt, 1e-6), 1e6)
        target = rate * args.cpu_seconds
        stride = max(1, int(np.ceil(np.sum(cost) / target)))
        sel = order[::stride]
        v, dt = run(sel)
        agree = flk.certificate_from_lk(pairs[sel], v) == flk.certificate_from_lk(pairs[sel], res.lk[sel])
        cpu = {
            "value": round(float(np.sum(cost[sel])) / dt * 1e-9, 4),
            "unit": "Gseg-pairs/s",
            "cores": int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)),
            "kind": "port",
            "sample": f"every {stride}th candidate pair ({len(sel)} pairs, "
                      f"{int(np.sum(cost[sel]))} segment pairs) through the oracle's restatement "
                      f"of the reference direct-sum pair loop (DirectFast per pair, OpenMP over "
                      f"pairs), {dt:.1f} s",
            "certificate_agrees": bool(agree),
        }
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
            "curves": len(curves),
            "vertices": int(offs[-1]),
            "candidate_pairs": int(len(pairs)),
            "seg_pairs": seg_pairs,
            "linked_pairs": len(cert),
            "certificate_sha1": hashlib.sha1(repr(cert).encode()).hexdigest(),
            "max_off_integer": float(off_int.max()) if len(off_int) else 0.0,
            "fallbacks": int(st["fallbacks"]),
            "direct_threshold": args.direct_threshold,
            "ds_seg_pairs": int(st["ds_seg_pairs"]),
            "bh_visits": int(st["bh_visits"]),
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_bh": learned,
    }
    eng.close()
    print(json.dumps(out), flush=True)


BH_BYTES_PER_VISIT = 336  # two 160-byte node records + 16 bytes of queue traffic


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
    FP32 direct sum (a 1/16 row sample scaled by 16 unless --full-ds: the
    direct sum's cost is exactly proportional to rows), Barnes–Hut with the
    reference's beta criterion (adaptive beta as the reference), and
    Barnes–Hut with the learned criterion."""
    from fastlinkingnumbers_amd import criterion
    if not os.path.exists(criterion.DEFAULT_WEIGHTS):
        log("no learned-criterion weights; skipping the learned comparison")
        return None
    curves = workloads.workload(5)
    a, b = curves
    res = {"workload": "random_walk_pair 2x8M (rg 200, fp32)", "vertices": [len(a), len(b)]}
    if args.full_ds:
        eng.set_curves(curves)
        r = eng.linking_numbers([[0, 1]], method="directsum", fallback_tol=0.0)
        res["ds_ms"] = round(r.stats["ms_total"], 1)
        res["ds_lk"] = float(r.lk[0])
        res["ds_timing"] = "full"
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
        eng.linking_numbers([[0, 1]], method="barneshut", **kw)  # warm-up
        r = eng.linking_numbers([[0, 1]], method="barneshut", **kw)
        st = r.stats
        dev = eng.linking_numbers([[0, 1]], method="barneshut", timing=1, **kw).stats  # kernel times
        res[name] = {"ms": round(st["ms_total"], 1), "lk": float(r.lk[0]),
                     "off_integer": float(abs(r.lk[0] - round(r.lk[0]))),
                     "visits": int(st["bh_visits"]), "far": int(st["bh_far"]),
                     "leaf_pairs": int(st["bh_leaf_pairs"]), "retries": int(st["retries"]),
                     "fallbacks": int(st["fallbacks"]),
                     "ms_build": round(st["ms_bh_build"], 2), "ms_traverse": round(st["ms_bh_traverse"], 2),
                     "ms_level_kernels": round(dev["ms_bh_level"], 2), "ms_mlp_kernels": round(dev["ms_mlp"], 2)}
        log(f"workload 5 {name}: {res[name]}")
    res["learned_speedup_vs_ds"] = round(res["ds_ms"] / res["learned"]["ms"], 1)
    res["learned_speedup_vs_beta"] = round(res["beta"]["ms"] / res["learned"]["ms"], 2)
    res["same_integer"] = round(res["beta"]["lk"]) == round(res["learned"]["lk"])
    return res


if __name__ == "__main__":
    main()
