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
library on the stream the kernels run on.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gpu-inner-outer_amd", "python"))

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
    sweeps = int(max(1, min(50, (budget_s * 1000.0 - probe["abstraction_ms"]) / per_sweep_ms)))
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="unicycle_91x105x42x10")
    ap.add_argument("--sweep-all-pairs", action="store_true",
                    help="test every pair every iteration (the reference's work schedule)")
    ap.add_argument("--cpu-budget", type=float, default=15.0, help="seconds of reference CPU work for the baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    from gsc_amd.solver import Solver

    case = workload(args.workload)
    print(f"[bench] workload {case.name}: {case.state_count()} states x {case.input_count()} inputs", file=sys.stderr,
          flush=True)
    solver = Solver(case, time_kernels=True, sweep_all_pairs=args.sweep_all_pairs)
    P = solver.pair_count
    for w in range(args.warmup):
        solver.build_abstraction()
        s = solver.solve()
        print(f"[bench] warmup {w}: {s['iterations']} iterations, {s['message']}", file=sys.stderr, flush=True)

    abs_ms, solve_ms, sat_ms, prefix_ms, red_ms, iters, tested, counted, slots = 0.0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        a = solver.build_abstraction()
        s = solver.solve()
        abs_ms += a
        solve_ms += s["solve_ms"]
        sat_ms += s["satisfaction_ms"]
        prefix_ms += s["prefix_ms"]
        red_ms += s["reduction_ms"]
        iters += s["iterations"]
        tested += s["pairs_tested"]
        counted += s["boxes_counted"]
        slots += s["test_slots"]
    wall_ms = (time.perf_counter() - t0) * 1000.0
    summary = s
    K = args.steps
    gpairs = P * iters / (sat_ms / 1000.0) / 1e9
    # Dominant kernel per step decides the roofline line.
    sat_launch_ms = sat_ms / iters
    S = solver.state_count
    sat_bytes_per_launch = (4.0 * tested + 7.0 * S * iters) / iters
    abs_launch_ms = abs_ms / K
    if sat_ms >= abs_ms:
        roof = {"kernel": "satisfy_kernel", "bound": "hbm", "achieved": sat_bytes_per_launch / (sat_launch_ms / 1000.0) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                "launch_ms": sat_launch_ms, "algorithmic_bytes_per_launch": sat_bytes_per_launch}
    else:
        abs_bytes = 4.0 * P
        roof = {"kernel": "abstraction_kernel", "bound": "hbm", "achieved": abs_bytes / (abs_launch_ms / 1000.0) / 1e9,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "traffic": None,
                "launch_ms": abs_launch_ms, "algorithmic_bytes_per_launch": abs_bytes}
    roof["frac"] = roof["achieved"] / roof["peak"]

    line = {
        "metric": "pair-satisfaction Gpairs/s (pairs x iterations / satisfaction time) and solve ms at 4.0M states x 100 inputs",
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
        },
        "abstraction_ms": abs_ms / K,
        "solve_ms": solve_ms / K,
        "total_ms": (abs_ms + solve_ms) / K,
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
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(case, args.cpu_budget)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
