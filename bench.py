"""Benchmark: continuous Frechet distance of one 3D pair of 200,000 x 200,000
points on one MI355X.

A step is one full distance computation (DFD upper bound, VE-FD lower bound,
bisection to 1e-9 relative) with both curves already resident in HBM.  The
value is cells/s: grid cells swept by all sweeps of the timed steps / wall
time.  Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64|mixed]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cells/s at 200k×200k pts; batched pairs/s learned vs plain bisection @1e-9 rel"

# Algorithmic FP64/FP32 flops per cell (add, sub, mul, div, sqrt = 1; FMA = 2;
# per-segment terms excluded).
FLOPS_PER_CELL = {"vefd_f64": lambda D: 16 * D - 2, "vefd_f32": lambda D: 16 * D - 2,
                  "decide_f64": lambda D: 10 * D + 18, "decide_f32": lambda D: 10 * D + 18,
                  "dfd": lambda D: 3 * D - 1}
PEAK_TFLOPS = {"vefd_f64": 78.6, "decide_f64": 78.6, "dfd": 78.6, "bounds_f64": 78.6,
               "vefd_f32": 157.3, "decide_f32": 157.3}
# Instantiation run for each kind on the session path: band_sweep<OP, dims, type,
# scoped hand-off, rows per lane, no pruning>.
_R = int(os.environ.get("RFD_ROWS", "4"))
KERNEL_SYMBOL = {k: f"rfd::band_sweep<{op}, 3, {t}, 1, {_R}, false>" for k, op, t in
                 (("vefd_f64", 1, "double"), ("vefd_f32", 1, "float"), ("decide_f64", 2, "double"),
                  ("decide_f32", 2, "float"), ("dfd", 0, "double"))}
# Fused DFD + VE-FD sweep: bounds_sweep<dims, rows per lane, producer waves,
# columns per step, mode (3: FP64)>; at 200k rows the library picks 13 rows
# per lane, 10 producer waves and one column per step.
_BCFG = int(os.environ.get("RFD_BOUNDS", "1310"))
_BC, _BBASE = (2, _BCFG - 20000) if _BCFG >= 20000 else (1, _BCFG)
_BR, _BN = (_BBASE // 100, _BBASE % 100) if _BBASE >= 1000 else (_BBASE // 10, _BBASE % 10)
KERNEL_SYMBOL["bounds_f64"] = f"rfd::bounds_sweep<3, {_BR}, {_BN}, {_BC}, 3>"


def flops_per_launch(kind: str, dims: int, m: int, n: int, evaluated_per_launch: float) -> float:
    """Algorithmic flops of one launch: the fused sweep does one DFD cell per
    point pair and one VE-FD cell per segment pair."""
    if kind == "bounds_f64":
        return FLOPS_PER_CELL["dfd"](dims) * m * n + FLOPS_PER_CELL["vefd_f64"](dims) * (m - 1) * (n - 1)
    return FLOPS_PER_CELL[kind](dims) * evaluated_per_launch


def sweep_mix(stats: dict, m: int, n: int) -> dict:
    """Per-sweep cell counts, the fused sweep counted as its DFD and VE-FD parts."""
    mix = {k: dict(v) for k, v in stats.items()}
    b = mix.pop("bounds_f64", None)
    if b:
        for k, c in (("dfd", m * n), ("vefd_f64", (m - 1) * (n - 1))):
            e = mix.setdefault(k, {"launches": 0, "ms": 0.0, "cells": 0.0, "evaluated": 0.0})
            e["cells"] += c * b["launches"]
    return mix


def _profiled(m: int, n: int, kernel: str):
    """The dominant kernel's entry of the committed rocprofv3 counter runs
    (profiles/hbm_traffic_band_sweep_200k.json), when they were taken on this
    workload; None otherwise."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_band_sweep_200k.json")
    if (m, n) != (200_000, 200_000) or not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    tag = {"vefd_f64": "band_sweep<1, 3, double", "dfd": "band_sweep<0, 3, double",
           "bounds_f64": KERNEL_SYMBOL["bounds_f64"].replace("rfd::", "")}.get(kernel)
    for name, v in data["kernels"].items():
        if tag and tag in name:
            return v
    return None


def measured_traffic(m: int, n: int, kernel: str):
    """HBM bytes per launch of the dominant kernel from the committed FETCH_SIZE
    / WRITE_SIZE runs, scaled by the copy-kernel calibration; None if absent."""
    v = _profiled(m, n, kernel)
    return v["traffic_bytes"] if v else None


def measured_valu_issue(m: int, n: int, kernel: str):
    """Fraction of the busy CUs' VALU issue cycles the kernel used, from the
    committed SQ_INSTS_VALU / GRBM_GUI_ACTIVE run (4 cycles per wave64 VALU
    instruction); None if absent."""
    v = _profiled(m, n, kernel)
    return v.get("valu_issue_frac") if v else None


def cpu_threads() -> int:
    """Host threads for the multi-core CPU timing: the cores this process may
    run on, at most 16."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    return max(1, min(16, avail))


def cpu_model() -> str:
    """The host CPU's model name as /proc/cpuinfo gives it."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _time_rows(run, rows0: int, rows_max: int, target_s: float, step: int = 1):
    """Grows a row sample until one run takes about target_s; returns
    (rows, seconds)."""
    rows = max(step, rows0 - rows0 % step)
    while True:
        t0 = time.perf_counter()
        run(rows)
        dt = time.perf_counter() - t0
        if dt > target_s or rows >= rows_max:
            return rows, dt
        grow = max(2.0, min(8.0, target_s / max(dt, 1e-4)))
        rows = min(rows_max, int(rows * grow))
        rows = max(step, rows - rows % step) if rows < rows_max else rows


def cpu_baseline(P, Q, mix: dict, dims: int, budget_s: float = 20.0) -> dict:
    """Times the CPU oracle on row samples of each sweep the GPU ran, on one
    thread and on the host's cores (row blocks pipelined over column chunks,
    bit-identical results), and prices the GPU run's per-kernel cell counts
    at those rates.  The multi-core rate is the reported value; the decision
    sweep, which the multi-core oracle does not split, is priced at its
    single-thread rate times the thread count measured on the VE-FD sweep."""
    from oracle import oracle as O

    n = len(Q)
    threads = cpu_threads()
    kinds = {"vefd_f64": 0, "vefd_f32": 0, "decide_f64": 1, "decide_f32": 1, "dfd": 2}
    used = sorted({kinds[k] for k in mix if k in kinds})
    per = budget_s / max(1, 2 * len(used))
    lb = O.endpoint_lb(P, Q)
    rate1, rate_mt, samples = {}, {}, []
    for kind in used:
        rmax = len(P) - 1 if kind != 2 else len(P)
        cols = n - 1 if kind != 2 else n
        rows, dt = _time_rows(lambda r: O.sweep_rows(P, Q, kind, lb * 1.5, 0, r), 8, rmax, per / 4)
        rate1[kind] = rows * cols / dt
        name = ["vefd", "decide", "dfd"][kind]
        samples.append(f"{name}: rows 0..{rows} x {cols} cols on 1 thread")
        if kind != 1 and threads > 1:
            rows, dt = _time_rows(lambda r: O.sweep_rows_mt(P, Q, kind, lb * 1.5, 0, r, threads),
                                  8 * threads, rmax, per / 4, step=threads)
            rate_mt[kind] = rows * cols / dt
            samples.append(f"{name}: rows 0..{rows} x {cols} cols on {threads} threads")
    if 1 in used:
        speed = rate_mt.get(0, rate1.get(0, 1.0)) / rate1.get(0, 1.0) if 0 in rate1 else threads
        rate_mt[1] = rate1[1] * speed
    for k in rate1:
        rate_mt.setdefault(k, rate1[k])
    total = sum(mix[k]["cells"] for k in mix if k in kinds)
    t1 = sum(mix[k]["cells"] / rate1[kinds[k]] for k in mix if k in kinds)
    tm = sum(mix[k]["cells"] / rate_mt[kinds[k]] for k in mix if k in kinds)
    return {"value": total / tm, "unit": "cells/s", "cores": threads, "kind": "port",
            "single_thread_value": total / t1, "cpu_model": cpu_model(),
            "host_cpus": os.cpu_count(),
            "sample": "CPU oracle (g++ -O3, FP64) on " + "; ".join(samples) +
                      "; rates applied to the GPU run's per-sweep cell counts"}


def precision_form(sess, ctx, ref: dict) -> dict:
    """One distance computation of the same resident pair in mixed precision:
    FP32 VE-FD weights and FP32 bisection while the bracket is wide, FP64 DFD,
    FP64 certification of the FP32 bracket and FP64 final steps.  Reported
    beside the FP64 line with its error against the FP64 result."""
    ctx.reset_stats()
    t0 = time.perf_counter()
    r = sess.distance(stop="rel", eps=1e-9, precision="mixed")
    wall = time.perf_counter() - t0
    st = ctx.stats()
    cells = sum(v["cells"] for v in st.values())
    return {"distance": r["distance"], "fp64_distance": ref["distance"],
            "rel_error_vs_fp64": abs(r["distance"] - ref["distance"]) / ref["distance"],
            "err_bound": r["err_bound"], "iterations": r["iterations"],
            "iterations_fp32": r["iterations_fp32"], "certify_steps": r["certify_steps"],
            "ms": wall * 1e3, "cells_per_s": cells / wall,
            "kernels": {k: dict(v, avg_ms=v["ms"] / v["launches"]) for k, v in st.items()}}


def robot_section(ctx) -> dict:
    """The long noisy-copy pair (3D robot-like paths of 100,000 and 120,000
    points, seed 0), in FP64 to 1e-9 relative.  Its bracket collapses (VE-FD
    equals DFD), so the run is the fused bound sweep and no decision."""
    from robotfd_amd import Intersection, workloads as W
    P, Q = W.robot_pair(100_000, 120_000, seed=0)
    with Intersection(P, Q, ctx=ctx) as s:
        s.data_move()
        ctx.reset_stats()
        t0 = time.perf_counter()
        r = s.distance(stop="rel", eps=1e-9)
        wall = time.perf_counter() - t0
    st = ctx.stats()
    cells = sum(v["cells"] for v in st.values())
    return {"workload": "3D robot-like pair 100,000 x 120,000 (noisy copy), seed 0; FP64 to 1e-9 relative",
            "distance": r["distance"], "dfd": r["dfd"], "vefd": r["vefd"], "iterations": r["iterations"],
            "ms": wall * 1e3, "cells_per_s": cells / wall,
            "kernels": {k: dict(v, avg_ms=v["ms"] / v["launches"]) for k, v in st.items()}}


def refined_ub_section(pairs, open_mask, ctx, levels=(1, 2)) -> dict:
    """What the tighter upper bound (DFD of the midpoint-refined curves) saves
    on the pairs whose bracket stays open: decision sweeps and time of their
    single-pair runs from [VE-FD, DFD] and from the refined bracket."""
    import numpy as np
    from robotfd_amd import Intersection
    idx = np.flatnonzero(open_mask)
    out = {"open_pairs": int(idx.size)}
    if not idx.size:
        return out
    runs = {0: [0, 0.0]}
    for lv in levels:
        runs[lv] = [0, 0.0]
    for k in idx:
        P, Q = pairs[k]
        with Intersection(P, Q, ctx=ctx) as s:
            for lv in runs:
                r = s.distance(stop="rel", eps=1e-9, refine_ub=lv)
                runs[lv][0] += int(r["iterations"])
                runs[lv][1] += r["ms_total"]
    out["decision_sweeps_plain"] = runs[0][0]
    out["ms_plain"] = runs[0][1]
    for lv in levels:
        out[f"levels_{lv}"] = {"decision_sweeps": runs[lv][0], "sweeps_saved": runs[0][0] - runs[lv][0],
                               "ms": runs[lv][1]}
    return out


def batch_section(count: int, ctx, refined: bool = True) -> dict:
    """Pairs/s of the batched pipeline on `count` pairs (lengths log-uniform in
    [500, 20000] per curve, dims 2/3 at random, N(0,1) random walks, seed 0;
    the network never saw these pairs): plain sequential bisection,
    speculative bisection (same results), and the learned bracket with either
    (results within 1e-9 relative of plain)."""
    import numpy as np
    from robotfd_amd import workloads as W
    from robotfd_amd.batch import Batch
    from robotfd_amd.learned import DEFAULT_WEIGHTS, BracketNet, run_learned

    pairs = W.batch_pairs(count, seed=0)
    res = {"workload": f"{count} pairs, lengths log-uniform [500, 20000] per curve, "
                       "dims 2/3 50/50, N(0,1) random walks, seed 0; to 1e-9 relative"}
    with Batch(pairs, ctx=ctx) as b:
        best, tune_ms = b.tune()
        res["rows_per_lane"] = best
        res["tune_ms"] = tune_ms
        ref = None
        net = None
        if os.path.exists(DEFAULT_WEIGHTS):
            net = BracketNet(ctx=ctx, path=DEFAULT_WEIGHTS)
        runs = [("plain", 1, None), ("speculative", 0, None)]
        if net is not None:
            runs += [("learned", 1, net), ("learned_speculative", 0, net)]
        for name, depth, nn in runs:
            go = (lambda: run_learned(b, nn, spec_depth=depth)) if nn else (lambda: b.run(spec_depth=depth))
            go()  # warm (graph capture)
            t0 = time.perf_counter()
            r = go()
            dt = time.perf_counter() - t0
            if ref is None:
                ref = r
            lat = r["latency_ms"]
            open_pairs = ref["iterations"] > 0
            entry = {"pairs_per_s": count / dt, "ms": dt * 1e3,
                     "p50_latency_ms": float(np.percentile(lat, 50)),
                     "p99_latency_ms": float(np.percentile(lat, 99)),
                     "p99_9_latency_ms": float(np.percentile(lat, 99.9)),
                     "max_latency_ms": float(lat.max()),
                     # Pairs whose bracket stays open after the bounds finish in
                     # the decision rounds: their latencies are the tail.
                     "open_pair_p50_latency_ms": float(np.percentile(lat[open_pairs], 50)) if open_pairs.any() else None,
                     "open_pair_max_latency_ms": float(lat[open_pairs].max()) if open_pairs.any() else None,
                     "decision_sweeps": int(r["iterations"].sum()),
                     "max_rel_diff_vs_plain": float(np.max(np.abs(r["distance"] - ref["distance"]) /
                                                           np.maximum(ref["distance"], 1e-300)))}
            if nn is not None:
                entry["certified_frac_of_open"] = float(r["certified"][open_pairs].mean()) if open_pairs.any() else 0.0
                entry["fallback_frac_of_open"] = 1.0 - entry["certified_frac_of_open"]
            else:
                entry["bisected_pairs"] = int(open_pairs.sum())
            res[name] = entry
        for name in ("speculative", "learned", "learned_speculative"):
            if name in res:
                res[name]["speedup_vs_plain"] = res[name]["pairs_per_s"] / res["plain"]["pairs_per_s"]
        # The bound round (DFD and VE-FD of every pair) ends where the closed
        # pairs finish; its share of the plain run caps what any decision-side
        # speed-up (speculation, the learned bracket) can gain.
        closed = ref["iterations"] == 0
        if closed.any():
            bound_ms = float(ref["latency_ms"][closed].max())
            res["bound_round_ms"] = bound_ms
            res["bound_round_frac_of_plain"] = bound_ms / res["plain"]["ms"]
            res["speedup_ceiling_if_decisions_free"] = res["plain"]["ms"] / bound_ms
        if refined:
            res["refined_ub"] = refined_ub_section(pairs, ref["iterations"] > 0, ctx)
        if net is not None:
            res["learned_policy"] = {"tau": net.policy[0], "delta": net.policy[1],
                                     "weights": os.path.relpath(DEFAULT_WEIGHTS, ROOT)}
            net.close()
    return res


def profile_mode(what: str, points: int) -> int:
    """One launch of the named kernel (after the bound sweep it needs), for
    counter runs under rocprofv3; prints what ran."""
    if what == "copy":
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nbytes = 256 << 20
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
        assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
        for _ in range(3):
            assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # hipMemcpyDeviceToDevice
        assert hip.hipDeviceSynchronize() == 0
        print(json.dumps({"profile": "copy", "bytes_read": nbytes, "bytes_written": nbytes, "copies": 3}))
        return 0
    from robotfd_amd import Context, Intersection, workloads as W
    if what == "learned":
        # One learned-bracket run of the 4,096-pair batch (its network layers
        # are the MFMA kernels).
        from robotfd_amd.batch import Batch
        from robotfd_amd.learned import DEFAULT_WEIGHTS, BracketNet, run_learned
        ctx = Context.default()
        with Batch(W.batch_pairs(4096, seed=0), ctx=ctx) as b:
            net = BracketNet(ctx=ctx, path=DEFAULT_WEIGHTS)
            try:
                r = run_learned(b, net, spec_depth=0)
            finally:
                net.close()
        print(json.dumps({"profile": what, "pairs": 4096, "iterations": int(r["iterations"].sum()),
                          "kernels": ctx.stats()}))
        return 0
    P, Q = W.long_3d(points, points, seed=0)
    ctx = Context.default()
    with Intersection(P, Q, ctx=ctx) as s:
        # "mixed": the bound sweep with the VE-FD in FP32; "dfd": its DFD-only
        # form; otherwise the FP64 bound sweep.
        dfd, ve = s.Bound_call_gpu(0.0, run_vefd=what != "dfd", fp32=what == "mixed")
        out = {"profile": what, "points": points, "dfd": dfd, "vefd": ve}
        if what == "decision":
            out["radius"] = dfd * (1 - 0.5e-9)
            out["terminal_ub"] = s.decide_ub(out["radius"])
        out["kernels"] = ctx.stats()
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precision", choices=["fp64", "mixed"], default="fp64")
    ap.add_argument("--points", type=int, default=200_000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-robot", action="store_true",
                    help="skip the 100k x 120k noisy-copy pair reported beside the main line")
    ap.add_argument("--no-precision-form", action="store_true",
                    help="skip the mixed-precision run reported beside the FP64 line")
    ap.add_argument("--batch-pairs", type=int, default=4096,
                    help="pairs of the batched workload reported beside the main line (0: skip)")
    ap.add_argument("--profile", choices=["decision", "bounds", "mixed", "dfd", "copy", "learned"],
                    help="run one launch of a kernel for rocprofv3 and exit: the FP64 decision at "
                         "DFD (1 - 0.5e-9) or the fused bound sweep on the pair of --points (FP64, "
                         "VE-FD in FP32, or DFD alone), or a "
                         "256 MiB device-to-device copy (byte-counter calibration), or one "
                         "learned-bracket run of the 4,096-pair batch")
    args = ap.parse_args()
    if args.profile:
        return profile_mode(args.profile, args.points)
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the path has no "
              "multi-GPU sharding in this build", file=sys.stderr)
        return 2

    from robotfd_amd import Context, Intersection, workloads as W

    m = n = args.points
    P, Q = W.long_3d(m, n, seed=0)
    sha = [W.sha256(P), W.sha256(Q)]
    with open(os.path.join(ROOT, "tests", "golden", "workload_sha256.json")) as f:
        want = json.load(f).get(f"long_3d_{m}x{n}_seed0")
    dims = 3
    ctx = Context.default()
    sess = Intersection(P, Q, ctx=ctx)
    sess.data_move()
    kw = dict(stop="rel", eps=1e-9, precision=args.precision)
    for w in range(args.warmup):
        r = sess.distance(**kw)
        print(f"warmup {w}: distance={r['distance']:.12f} iterations={r['iterations']} "
              f"ms={r['ms_total']:.1f}", file=sys.stderr, flush=True)
    ctx.reset_stats()
    results = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        results.append(sess.distance(**kw))
        print(f"step {k}: ms={results[-1]['ms_total']:.1f}", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    stats = ctx.stats()
    wall = t1 - t0
    cells = sum(v["cells"] for v in stats.values())
    value = cells / wall

    dom = max((k for k in stats if k in PEAK_TFLOPS), key=lambda k: stats[k]["ms"])
    st = stats[dom]
    avg_ms = st["ms"] / st["launches"]
    # Flops of the cells the kernel actually evaluated (pruned / unreachable
    # cells are skipped and not counted).
    flops = flops_per_launch(dom, dims, m, n, st["evaluated"] / st["launches"])
    achieved = flops / (avg_ms * 1e-3) / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS[dom], "unit": "TFLOP/s",
            "frac": achieved / PEAK_TFLOPS[dom], "traffic": measured_traffic(m, n, dom),
            "kernel": KERNEL_SYMBOL[dom], "avg_launch_ms": avg_ms,
            "flops_per_launch": flops,
            "cells_per_launch": st["cells"] / st["launches"],
            "evaluated_cells_per_launch": st["evaluated"] / st["launches"]}
    if roof["traffic"]:
        roof["hbm_gbs"] = roof["traffic"] / (avg_ms * 1e-3) / 1e9  # profiled bytes over this run's launch time
    roof["valu_issue_frac"] = measured_valu_issue(m, n, dom)

    out = {
        "metric": METRIC, "value": value, "unit": "cells/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": wall * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"3D random walks {m:,} x {n:,} points, N(0,1) steps, seed 0; "
                               f"DFD + VE-FD + bisection to 1e-9 relative, {args.precision}",
                   "points": [m, n], "dims": dims, "stop": "rel 1e-9", "precision": args.precision,
                   "inputs_sha256": sha, "inputs_match_fixture": (sha == want) if want else None},
        "roofline": roof,
        "result": {"distance": results[-1]["distance"], "iterations": results[-1]["iterations"],
                   "dfd": results[-1]["dfd"], "vefd": results[-1]["vefd"],
                   "err_bound": results[-1]["err_bound"]},
        "kernels": {k: dict(v, avg_ms=v["ms"] / v["launches"],
                            cells_per_s=v["cells"] / (v["ms"] * 1e-3)) for k, v in stats.items()},
    }
    if args.precision == "fp64" and not args.no_precision_form:
        out["precision_form"] = precision_form(sess, ctx, results[-1])
    sess.free_memory()
    if not args.no_robot:
        out["robot_pair"] = robot_section(ctx)
    if args.batch_pairs:
        out["batch"] = batch_section(args.batch_pairs, ctx)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(P, Q, sweep_mix(stats, m, n), dims)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
