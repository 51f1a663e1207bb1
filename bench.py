"""Benchmark of the grid symbolic-control hot path on one MI355X.

A step is one complete solve of the workload with its case already resident
in HBM: the pair abstraction (successor cell boxes of every state-input pair)
followed by the backward fixpoint to convergence. `value` is the
pair-satisfaction throughput in the reference's definition (README.md:66,
88-98): state-input pairs x iterations / total pair-satisfaction kernel time,
counting every pair in every iteration.

Default workload: the 91x105x42x10 unicycle grid x 100 inputs (401,310,000
pairs) with a hyperbolic target band. `--workload hyperbolic_demo` runs the
reference's published case instead.

Prints one JSON line. Kernel times come from HIP events recorded by the
library on the stream the kernels run on. `roofline` prices the dominant
kernel at 9 B (the reference's per-pair record: u32 min_flat + u32 max_flat +
u8 valid) per pair it processes; `roofline.traffic` is that kernel's measured
HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, calibrated on a plain
copy); `cpu_baseline` is the reference's own CPU build on a bounded slice of
the same workload, on one host core.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-inner-outer_amd", "python"))

METRIC = "solve ms & Gpairs/s at 4.0M states×100 inputs; served solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0
REF_PUBLISHED_GPAIRS = 23.2  # RTX 4090, hyperbolic_demo, README.md:66 / 88-98


def workload(name):
    from gsc_amd import cases

    if name == "unicycle_91x105x42x10":
        return cases.unicycle_91x105x42x10()
    if name == "hyperbolic_demo":
        return cases.hyperbolic_demo()
    if name == "unicycle_128x128x64x16":
        return cases.unicycle_128x128x64x16()
    if name == "unicycle_30x30x24x10":
        return cases.unicycle_30x30x24x10(sample_time=0.5)
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(case, budget_s):
    """The reference CPU build (oracle/_ref) timed on a bounded slice of the same
    workload: abstraction of the slice, then sweeps of the pair-satisfaction test
    over the slice's pairs (two prefix queries per valid pair, the reference
    kernel's per-pair work)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import gsc_oracle

    if not gsc_oracle.ref_available():
        return None
    S = case.state_count()
    I = case.input_count()
    # States in the middle of the grid (valid, non-trivial boxes); ~2M pairs.
    n_states = max(1, min(S, 2_000_000 // I))
    begin = max(0, S // 2 - n_states // 2)
    probe = gsc_oracle.ref_sample(case, begin, begin + n_states, 1)
    per_sweep_ms = max(probe["satisfaction_ms"], 1e-3)
    sweeps = int(max(1, min(400, (budget_s * 1000.0 - probe["abstraction_ms"]) / per_sweep_ms)))
    r = gsc_oracle.ref_sample(case, begin, begin + n_states, sweeps)
    pairs = r["sample_pairs"]
    return {
        "value": pairs * sweeps / (r["satisfaction_ms"] / 1000.0) / 1e9,
        "unit": "Gpairs/s",
        "cores": 1,
        "kind": "reference",
        "sample": (f"reference CPU build (-O3, 1 thread): {int(pairs)} pairs = states [{begin}, {begin + n_states}) "
                   f"x {I} inputs of the same workload; pair-satisfaction test swept {sweeps}x over them "
                   f"({r['satisfaction_ms'] / 1000:.1f} s); abstraction of the slice "
                   f"{r['abstraction_ms'] / 1000:.2f} s = {pairs / (r['abstraction_ms'] / 1000.0) / 1e6:.2f} Mpairs/s"),
        "abstraction_mpairs_per_s": pairs / (r["abstraction_ms"] / 1000.0) / 1e6,
        "cpu": _cpu_model(),
    }


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# ---------------------------------------------------------------------------
# HBM traffic of the dominant kernel, from rocprofv3 hardware counters.
#
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch, summed over the TCC channels)
# are collected in two separate rocprofv3 passes over a one-step child run of
# this same script, and each is scaled by what the same counter reads for a
# plain 1 GiB device-to-device copy (the runtime's copy kernel) measured in
# the same way, so that the reported figure is in bytes of HBM traffic.
# ---------------------------------------------------------------------------
COPY_BYTES = 1 << 30
COPY_REPEATS = 3


def copy_calibration():
    """Child: three 1 GiB device-to-device copies and nothing else."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (src, dst):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(COPY_BYTES)) != 0:
            raise SystemExit("hipMalloc failed")
    for _ in range(COPY_REPEATS):
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(COPY_BYTES), 3) != 0:  # hipMemcpyDeviceToDevice
            raise SystemExit("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)
    print(json.dumps({"copied_bytes": COPY_BYTES, "repeats": COPY_REPEATS}))


def _pmc_pass(counter, child_args, outdir, timeout_s=150):
    import csv
    import subprocess

    d = os.path.join(outdir, counter + "_" + ("copy" if "--copy-calibration" in child_args else "run"))
    cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    print(f"[bench] pmc pass {counter}: {' '.join(child_args)}", file=sys.stderr, flush=True)
    p = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout_s)
    if p.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counter} exited {p.returncode}: {p.stderr[-400:]}")
    child = None
    for line in p.stdout.splitlines():
        if line.startswith("{"):
            child = json.loads(line)
    per_kernel = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] != counter:
                continue
            k = row["Kernel_Name"].split("(")[0].split("<")[0].split("::")[-1]  # bare kernel name
            n, s = per_kernel.get(k, (0, 0.0))
            per_kernel[k] = (n + 1, s + float(row["Counter_Value"]))
    return child, per_kernel


def measure_traffic(workload_name, kernel, sweep_all, outdir):
    """HBM bytes per launch of `kernel`: (FETCH_SIZE x f_fetch + WRITE_SIZE x f_write) x 1024,
    f = known copy bytes / counter reading of the copy kernel, averaged over the work launches."""
    run_args = ["--workload", workload_name, "--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--traffic", "off",
                "--no-batching", "--no-serving"]
    if sweep_all:
        run_args.append("--sweep-all-pairs")
    out = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        _, copy_k = _pmc_pass(counter, ["--copy-calibration"], outdir)
        n_copy, kib_copy = copy_k.get("__amd_rocclr_copyBuffer", (0, 0.0))
        if kib_copy <= 0:
            raise RuntimeError(f"copy kernel not seen in the {counter} calibration pass")
        factor = (COPY_BYTES / 1024.0) * COPY_REPEATS / kib_copy
        child, run_k = _pmc_pass(counter, run_args, outdir)
        n, kib = run_k.get(kernel, (0, 0.0))
        if n == 0 or child is None:
            raise RuntimeError(f"{kernel} not seen in the {counter} pass")
        # The satisfaction kernel is launched in batches of 8; launches after the
        # fixpoint has stopped return at once. Average over the iterations run
        # (every solve of the child, untimed and timed).
        units = child["iterations_run"] if kernel == "satisfy_kernel" else n
        out[counter] = {"kib_total": kib, "dispatches": n, "launches_counted": units,
                        "copy_kernel_dispatches": n_copy, "copy_kib_per_gib": kib_copy / COPY_REPEATS,
                        "factor": factor, "bytes_per_launch": kib * 1024.0 * factor / units}
    out["bytes_per_launch"] = out["FETCH_SIZE"]["bytes_per_launch"] + out["WRITE_SIZE"]["bytes_per_launch"]
    return out


def _same_result(a, b):
    keys = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step")
    return all(np.array_equal(getattr(a, k), getattr(b, k)) for k in keys) and a.iterations == b.iterations \
        and a.result_message == b.result_message


def batching_section(count=32, seed=3, steps=3):
    """32 uneven reach-avoid cases: plain single-case solves one after another
    versus one batch (fused stages, captured graph, autotuned tiles). A step is
    abstraction + fixpoint of every case; setup (tables, masks, allocation) is
    outside both timings."""
    from gsc_amd import cases
    from gsc_amd.batch import Batch
    from gsc_amd.solver import Solver

    cs = cases.batch_instances(count, seed=seed)
    print(f"[bench] batching: {count} cases, {sum(c.state_count() * c.input_count() for c in cs)} pairs",
          file=sys.stderr, flush=True)
    solvers = [Solver(c, time_kernels=False) for c in cs]
    for s in solvers:
        s.build_abstraction()
        s.solve()
    seq = []
    for _ in range(steps):
        t0 = time.perf_counter()
        for s in solvers:
            s.build_abstraction()
            s.solve()
        seq.append((time.perf_counter() - t0) * 1000.0)
    b = Batch(cs)
    b.build_abstraction()
    t0 = time.perf_counter()
    tried = b.tune()
    tune_s = time.perf_counter() - t0
    bat, summ = [], None
    for _ in range(steps):
        t0 = time.perf_counter()
        b.build_abstraction()
        summ = b.solve()
        bat.append((time.perf_counter() - t0) * 1000.0)
    identical = all(_same_result(b.download(k), solvers[k].download()) for k in range(count))
    opts = b.options
    b.close()
    for s in solvers:
        s.close()
    seq_ms, bat_ms = sum(seq) / steps, sum(bat) / steps
    return {
        "workload": f"batch_instances(count={count}, seed={seed}): cells log-uniform in [5e4, 1e6], extents ~9:10:4:1, "
                    "inputs in {16, 36, 64, 100}, 1-3 targets, 0-5 obstacles",
        "cases": count,
        "pairs": int(sum(c.state_count() * c.input_count() for c in cs)),
        "batched_ms_per_step": bat_ms,
        "sequential_ms_per_step": seq_ms,
        "instances_per_s": count / (bat_ms / 1000.0),
        "sequential_instances_per_s": count / (seq_ms / 1000.0),
        "speedup": seq_ms / bat_ms,
        "graph_replays": summ["replays"],
        "max_iterations": summ["max_iterations"],
        "tuned": {k: opts[k] for k in ("tile_wg", "tile_ib", "iterations_per_launch", "order")},
        "tune_configs_tried": len(tried),
        "tune_s": tune_s,
        "identical_to_single_case_solves": identical,
    }


def serving_section(count=48, seed=5, rates=(100.0, 400.0, 1600.0)):
    """Poisson streams of solve requests (cells log-uniform in [2e4, 4e6]) at
    several rates: the continuous-batching server versus the plain per-call path
    serving the same stream one request at a time."""
    from gsc_amd import cases
    from gsc_amd.serve import Server, poisson_load, sequential_load

    reqs = cases.serving_requests(count, seed=seed)
    out = []
    identical = None
    for i, rate in enumerate(rates):
        keep = i == 0
        print(f"[bench] serving: {count} requests at {rate:g}/s", file=sys.stderr, flush=True)
        with Server(max_active=64, iterations_per_step=4, hbm_budget_bytes=96 << 30) as srv:
            served, r_srv = poisson_load(srv, reqs, rate, seed=int(rate), keep_results=keep)
        seq, r_seq = sequential_load(reqs, rate, seed=int(rate), keep_results=keep)
        if keep:
            identical = all(_same_result(a, b) for a, b in zip(r_srv, r_seq))
            del r_srv, r_seq
        out.append({
            "rate_per_s": rate,
            "served": {k: served[k] for k in ("solves_per_s", "p50_ms", "p99_ms", "mean_ms", "max_ms")},
            "server": {k: served["server"][k] for k in ("max_active_seen", "steps", "budget_peak")},
            "failed": len(served["failed"]),
            "sequential": {k: seq[k] for k in ("solves_per_s", "p50_ms", "p99_ms", "mean_ms", "max_ms")},
        })
    return {
        "workload": f"serving_requests(count={count}, seed={seed}): cells log-uniform in [2e4, 4e6], "
                    "inputs in {16, 36, 64, 100}, 1-3 targets, 0-5 obstacles; Poisson arrivals",
        "server": "max_active 64, 4 iterations per step, 96 GiB HBM budget",
        "rates": out,
        "identical_to_plain_path": identical,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="unicycle_91x105x42x10")
    ap.add_argument("--sweep-all-pairs", action="store_true",
                    help="test every pair every iteration (the reference's work schedule)")
    ap.add_argument("--recompute-abstraction", action="store_true",
                    help="store no pair records; recompute the tested successor boxes in FP64 every iteration")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of reference CPU work for the baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--traffic", choices=("live", "off"), default="live",
                    help="measure the dominant kernel's HBM bytes with two rocprofv3 --pmc passes")
    ap.add_argument("--pmc-dir", default=os.path.join(ROOT, "gpurun_out", "bench_pmc"))
    ap.add_argument("--copy-calibration", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-batching", action="store_true", help="skip the 32-case batching comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the Poisson serving comparison")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    from gsc_amd.solver import Solver

    case = workload(args.workload)
    print(f"[bench] workload {case.name}: {case.state_count()} states x {case.input_count()} inputs", file=sys.stderr,
          flush=True)
    K = args.steps
    iterations_run = 0  # satisfaction launches that did work, over both solvers (the traffic passes divide by it)
    # End to end: abstraction + fixpoint as a caller runs them, with no HIP
    # events between the kernels (each event record costs the stream a marker
    # packet and a gap of several microseconds).
    solver = Solver(case, time_kernels=False, sweep_all_pairs=args.sweep_all_pairs,
                    recompute_abstraction=args.recompute_abstraction)
    P = solver.pair_count
    for w in range(args.warmup):
        solver.build_abstraction()
        s = solver.solve()
        iterations_run += s["iterations"]
        print(f"[bench] warmup {w}: {s['iterations']} iterations, {s['message']}", file=sys.stderr, flush=True)
    abs_ms, solve_ms, iters = 0.0, 0.0, 0
    t0 = time.perf_counter()
    for k in range(K):
        abs_ms += solver.build_abstraction()
        s = solver.solve()
        solve_ms += s["solve_ms"]
        iters += s["iterations"]
    wall_ms = (time.perf_counter() - t0) * 1000.0
    iterations_run += iters
    summary = s
    solver.close()
    # Per-kernel breakdown: the same solves with HIP events around every
    # kernel of every iteration.
    solver = Solver(case, time_kernels=True, sweep_all_pairs=args.sweep_all_pairs,
                    recompute_abstraction=args.recompute_abstraction)
    solver.build_abstraction()
    iterations_run += solver.solve()["iterations"]
    sat_ms, prefix_ms, red_ms, timed_solve_ms, t_iters, tested, counted, slots = 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0
    for k in range(K):
        solver.build_abstraction()
        s = solver.solve()
        timed_solve_ms += s["solve_ms"]
        sat_ms += s["satisfaction_ms"]
        prefix_ms += s["prefix_ms"]
        red_ms += s["reduction_ms"]
        t_iters += s["iterations"]
        tested += s["pairs_tested"]
        counted += s["boxes_counted"]
        slots += s["test_slots"]
    iterations_run += t_iters
    if t_iters != iters:
        raise RuntimeError(f"timed solves ran {t_iters} iterations, untimed {iters}")
    gpairs = P * iters / (sat_ms / 1000.0) / 1e9
    # Dominant kernel per step decides the roofline line. Algorithmic bytes are
    # the reference pair record (u32 min_flat + u32 max_flat + u8 valid = 9 B)
    # per pair the launch processes: every pair in the reference's schedule,
    # the pairs of still-unsettled states in the default one.
    sat_launch_ms = sat_ms / iters
    S = solver.state_count
    abs_launch_ms = abs_ms / K
    if sat_ms >= abs_ms:
        units = tested / iters
        kname = "satisfy_recompute_kernel" if args.recompute_abstraction else "satisfy_kernel"
        roof = {"kernel": kname, "bound": "hbm", "launch_ms": sat_launch_ms,
                "algorithmic_bytes_per_launch": 9.0 * units, "pairs_processed_per_launch": units,
                "effective_GBs_all_pairs": 9.0 * P / (sat_launch_ms / 1000.0) / 1e9}
    else:
        roof = {"kernel": "abstraction_kernel", "bound": "hbm", "launch_ms": abs_launch_ms,
                "algorithmic_bytes_per_launch": 9.0 * P, "pairs_processed_per_launch": P}
    roof["achieved"] = roof["algorithmic_bytes_per_launch"] / (roof["launch_ms"] / 1000.0) / 1e9
    roof.update({"peak": HBM_PEAK_GBS, "unit": "GB/s"})
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None

    line = {
        "metric": METRIC,
        "value": gpairs,
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall_ms / K,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": case.name,
            "state_shape": list(solver.state_shape),
            "states": S,
            "inputs": solver.input_count,
            "pairs": P,
            "iterations": summary["iterations"],
            "stop": summary["message"],
            "schedule": "sweep_all_pairs" if args.sweep_all_pairs else "unsettled_states",
            "abstraction_dtype": "f64",
            "abstraction": "recomputed in every satisfaction launch" if args.recompute_abstraction
                           else "stored, 4 B per pair",
        },
        "abstraction_ms": abs_ms / K,
        "solve_ms": solve_ms / K,
        "total_ms": (abs_ms + solve_ms) / K,
        "solve_ms_with_kernel_events": timed_solve_ms / K,
        "iterations_run": iterations_run,
        "satisfaction_ms_per_iteration": sat_launch_ms,
        "prefix_ms_per_iteration": prefix_ms / iters,
        "bookkeeping_ms_per_iteration": red_ms / iters,
        "pairs_tested_per_iteration": tested / iters,
        "boxes_counted_per_iteration": counted / iters,
        "test_lane_utilisation": tested / max(1, slots),
        "reference_published_rtx4090_gpairs": REF_PUBLISHED_GPAIRS,
        "roofline": roof,
        "cpu_baseline": None,
    }
    solver.close()
    if args.traffic == "live" and not args.recompute_abstraction:
        try:
            t = measure_traffic(args.workload, roof["kernel"], args.sweep_all_pairs, os.path.abspath(args.pmc_dir))
            roof["traffic"] = t["bytes_per_launch"]
            roof["traffic_detail"] = t
        except Exception as e:  # counters unavailable: traffic stays null, the reason is kept
            roof["traffic_error"] = str(e)[:300]
    if not args.no_batching:
        line["batching"] = batching_section()
    if not args.no_serving:
        line["serving"] = serving_section()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_budget)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
