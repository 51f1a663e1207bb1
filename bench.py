"""Benchmark of the grid symbolic-control hot path on one MI355X.

A step is one complete solve of the workload with its case already resident
in HBM: the pair abstraction (host model tables, their upload, the table
kernels and the successor-box kernel) followed by the backward fixpoint to
convergence.

`value` is the pair-satisfaction throughput on the pairs the pass actually
tests: pairs tested per second of satisfaction kernel time (HIP events on the
library's stream). `all_pairs_priced_gpairs_per_s` prices every pair in every
iteration as the reference's definition does (README.md:66, 88-98); that
figure is only a work rate under the sweep schedule, which
`schedules_and_workloads.sweep_all_pairs` measures.

Default workload: the 91x105x42x10 unicycle grid x 100 inputs (401,310,000
pairs) with a hyperbolic target band. The line also carries the other
schedule (every pair every iteration), the reference's published case
(70x70x63x13 x 100), the 128x128x64x16 x 100 case with stored and recomputed
abstraction, the abstraction export, batching and serving.

`roofline` prices the dominant kernel at 9 B (the reference's per-pair
record: u32 min_flat + u32 max_flat + u8 valid) per pair it processes;
`roofline.traffic` is that kernel's measured HBM bytes per launch (the L2's
memory requests by size, two rocprofv3 --pmc passes); `cpu_baseline` is the
reference's own CPU build on a bounded slice of the same workload, on one
host core.
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
# The L2 (TCC) counts the requests it sends to memory by size: reads of 32,
# 64 and 128 bytes (TCC_EA0_RDREQ_32B / _64B / _128B), writes in total and of
# 64 bytes (TCC_EA0_WRREQ, TCC_EA0_WRREQ_64B; the rest are 32-byte writes).
# Bytes are those counts times their sizes, summed over the L2 channels: no
# calibration factor is involved. Two rocprofv3 --pmc passes (the hardware
# collects at most four TCC counters per pass) over a one-step child run of
# this script; the second also collects TCC_HIT / TCC_MISS. The same passes
# over three plain 1 GiB device-to-device copies (the runtime's copy kernel)
# must read back 1 GiB read and 1 GiB written per copy; their ratio is
# reported as a check of the counting.
# ---------------------------------------------------------------------------
COPY_BYTES = 1 << 30
COPY_REPEATS = 3
READ_PASS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")
WRITE_PASS = ("TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum", "TCC_HIT_sum", "TCC_MISS_sum")


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


def _pmc_pass(counters, child_args, outdir, timeout_s=150):
    """One rocprofv3 --pmc pass; returns (child's JSON line, {kernel: {counter: (dispatches, total)}})."""
    import csv
    import subprocess

    tag = ("copy" if "--copy-calibration" in child_args else "run") + "_" + counters[0].split("_sum")[0]
    d = os.path.join(outdir, tag)
    cmd = ["rocprofv3", "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--",
           sys.executable, os.path.abspath(__file__)] + child_args
    env = dict(os.environ, TMPDIR="/tmp")
    print(f"[bench] pmc pass {' '.join(counters)}: {' '.join(child_args)}", file=sys.stderr, flush=True)
    p = subprocess.run(cmd, cwd="/tmp", env=env, capture_output=True, text=True, timeout=timeout_s)
    if p.returncode != 0:
        raise RuntimeError(f"rocprofv3 --pmc {counters} exited {p.returncode}: {p.stderr[-400:]}")
    child = None
    for line in p.stdout.splitlines():
        if line.startswith("{"):
            child = json.loads(line)
    per_kernel = {}
    with open(os.path.join(d, "run_counter_collection.csv")) as f:
        for row in csv.DictReader(f):
            if row["Counter_Name"] not in counters:
                continue
            k = row["Kernel_Name"].split("(")[0].split("<")[0].split("::")[-1]  # bare kernel name
            n, s = per_kernel.setdefault(k, {}).get(row["Counter_Name"], (0, 0.0))
            per_kernel[k][row["Counter_Name"]] = (n + 1, s + float(row["Counter_Value"]))
    return child, per_kernel


def _bytes(c):
    rd = 32.0 * c["TCC_EA0_RDREQ_32B_sum"] + 64.0 * c["TCC_EA0_RDREQ_64B_sum"] + 128.0 * c["TCC_EA0_RDREQ_128B_sum"]
    w64 = c["TCC_EA0_WRREQ_64B_sum"]
    wr = 64.0 * w64 + 32.0 * (c["TCC_EA0_WRREQ_sum"] - w64)
    return rd, wr


def measure_traffic(workload_name, kernel, sweep_all, outdir):
    """HBM bytes per launch of `kernel` from the L2's memory requests by size,
    averaged over the launches that did work."""
    run_args = ["--workload", workload_name, "--steps", "1", "--warmup", "0", "--no-cpu-baseline", "--traffic", "off",
                "--no-batching", "--no-serving", "--no-extra"]
    if sweep_all:
        run_args.append("--sweep-all-pairs")
    totals, copy_totals, child, n_disp, n_copy = {}, {}, None, 0, 0
    for counters in (READ_PASS, WRITE_PASS):
        _, copy_k = _pmc_pass(counters, ["--copy-calibration"], outdir)
        ck = copy_k.get("__amd_rocclr_copyBuffer", {})
        if not ck:
            raise RuntimeError(f"copy kernel not seen in the {counters} pass")
        child, run_k = _pmc_pass(counters, run_args, outdir)
        kk = run_k.get(kernel, {})
        if not kk or child is None:
            raise RuntimeError(f"{kernel} not seen in the {counters} pass")
        for c in counters:
            totals[c] = kk[c][1]
            n_disp = kk[c][0]
            copy_totals[c] = ck[c][1]
            n_copy = ck[c][0]
    # The satisfaction kernel is launched in batches of 8; launches after the
    # fixpoint has stopped return at once. Average over the iterations run
    # (every solve of the child, untimed and timed).
    units = child["iterations_run"] if kernel == "satisfy_kernel" else n_disp
    rd, wr = _bytes(totals)
    crd, cwr = _bytes(copy_totals)
    known = float(COPY_BYTES) * COPY_REPEATS
    hits, misses = totals["TCC_HIT_sum"], totals["TCC_MISS_sum"]
    return {
        "read_bytes_per_launch": rd / units, "write_bytes_per_launch": wr / units,
        "bytes_per_launch": (rd + wr) / units, "launches_counted": units, "dispatches": n_disp,
        "read_requests": {k: totals[k] for k in READ_PASS}, "write_requests": {k: totals[k] for k in WRITE_PASS[:2]},
        "l2_hit_rate": hits / max(1.0, hits + misses), "l2_hits": hits, "l2_misses": misses,
        "copy_check": {"dispatches": n_copy, "read_measured_over_known": crd / known,
                       "write_measured_over_known": cwr / known},
        "method": "32/64/128-byte L2-to-memory read requests and 32/64-byte write requests times their sizes",
    }


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


def batch_direct_section(count=32, seed=3, steps=2):
    """The same 32 cases with the stage kernels launched directly instead of
    replayed from a captured graph (tile 256 x 8 inputs, 32 iterations per
    round, largest case first): the form profiled kernel by kernel."""
    from gsc_amd import cases
    from gsc_amd.batch import Batch

    cs = cases.batch_instances(count, seed=seed)
    with Batch(cs, tile_wg=1024, tile_ib=8, iterations_per_launch=32, order=1, use_graph=0) as b:
        b.build_abstraction()
        b.solve()
        ms = []
        for _ in range(steps):
            t0 = time.perf_counter()
            b.build_abstraction()
            summ = b.solve()
            ms.append((time.perf_counter() - t0) * 1000.0)
    return {"cases": count, "direct_ms_per_step": sum(ms) / steps, "rounds": summ["replays"],
            "max_iterations": summ["max_iterations"]}


def serving_section(count=48, seed=5, rates=(100.0, 400.0, 1600.0)):
    """Poisson streams of solve requests (cells log-uniform in [2e4, 4e6]) at
    several rates: the continuous-batching server versus the plain per-call path
    serving the same stream one request at a time."""
    from gsc_amd import cases
    from gsc_amd.serve import Server, poisson_load, sequential_load

    from gsc_amd.solver import run_case_u32

    reqs = cases.serving_requests(count, seed=seed)
    # Both paths warmed up before any timed stream: the server's kernels and the
    # per-call path's allocations of the largest requests.
    for c in sorted(reqs, key=lambda c: -c.state_count() * c.input_count())[:3]:
        run_case_u32(c)
    with Server(max_active=4, iterations_per_step=4, hbm_budget_bytes=8 << 30) as srv:
        poisson_load(srv, reqs[:4], 1e6, seed=0)
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


def _sha(a):
    import hashlib

    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _golden(name):
    path = os.path.join(ROOT, "tests", "golden", f"{name}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


GOLDEN_OF = {"unicycle_91x105x42x10": "unicycle_91x105x42x10", "hyperbolic_demo": "hyperbolic_demo"}
STATE_ARRAYS = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step")


def _matches_reference(res, fx):
    """Digests and per-iteration counts of a result against the reference build's golden."""
    if fx is None:
        return None
    return all(_sha(getattr(res, k)) == fx["sha256"][k] for k in STATE_ARRAYS) and \
        res.iterations == fx["iterations"] and res.result_message == fx["message"] and \
        [s["newly_reachable"] for s in res.iteration_stats] == fx["newly_reachable"]


def measure_workload(name, steps, warmup, sweep=False, recompute=False, log=True):
    """Abstraction + fixpoint of one workload, `steps` times, as a caller runs them
    (no per-kernel events), then the same solves with HIP events around every
    kernel for the per-kernel breakdown. Abstraction time is the wall time of
    gsc_solver_build_abstraction: host model tables, their upload, the table
    kernels and the abstraction kernel."""
    from gsc_amd.solver import Solver

    case = workload(name)
    solver = Solver(case, time_kernels=False, sweep_all_pairs=sweep, recompute_abstraction=recompute)
    P, S = solver.pair_count, solver.state_count
    iterations_run = 0
    for w in range(warmup):
        solver.build_abstraction()
        s = solver.solve()
        iterations_run += s["iterations"]
        if log:
            print(f"[bench] {name}{' sweep' if sweep else ''}{' recompute' if recompute else ''} warmup {w}: "
                  f"{s['iterations']} iterations, {s['message']}", file=sys.stderr, flush=True)
    abs_ms, abs_dev_ms, solve_ms, iters = 0.0, 0.0, 0.0, 0
    t0 = time.perf_counter()
    for _ in range(steps):
        ta = time.perf_counter()
        abs_dev_ms += solver.build_abstraction()
        abs_ms += (time.perf_counter() - ta) * 1000.0
        s = solver.solve()
        solve_ms += s["solve_ms"]
        iters += s["iterations"]
    wall_ms = (time.perf_counter() - t0) * 1000.0
    iterations_run += iters
    summary = s
    result = solver.download()
    solver.close()
    solver = Solver(case, time_kernels=True, sweep_all_pairs=sweep, recompute_abstraction=recompute)
    solver.build_abstraction()
    iterations_run += solver.solve()["iterations"]
    sat_ms, prefix_ms, red_ms, timed_solve_ms, t_iters, tested, counted, slots = 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0
    for _ in range(steps):
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
    solver.close()
    if t_iters != iters:
        raise RuntimeError(f"{name}: timed solves ran {t_iters} iterations, untimed {iters}")
    sat_launch_ms = sat_ms / iters
    out = {
        "workload": case.name, "state_shape": list(result.shape["state_shape"]), "states": S,
        "inputs": solver.input_count, "pairs": P,
        "schedule": "sweep_all_pairs" if sweep else "unsettled_states",
        "abstraction": "recomputed in every satisfaction launch" if recompute else "stored, 4 B per pair",
        "iterations": summary["iterations"], "stop": summary["message"],
        "abstraction_ms": abs_ms / steps, "abstraction_device_ms": abs_dev_ms / steps,
        "solve_ms": solve_ms / steps, "total_ms": (abs_ms + solve_ms) / steps, "ms_per_step": wall_ms / steps,
        "solve_ms_with_kernel_events": timed_solve_ms / steps,
        "satisfaction_ms_per_iteration": sat_launch_ms,
        "prefix_ms_per_iteration": prefix_ms / iters, "bookkeeping_ms_per_iteration": red_ms / iters,
        "pairs_tested_per_iteration": tested / iters, "boxes_counted_per_iteration": counted / iters,
        "test_lane_utilisation": tested / max(1, slots),
        # Pairs the satisfaction pass processed, per second of its kernel time.
        "processed_gpairs_per_s": tested / (sat_ms / 1000.0) / 1e9,
        # Every pair priced in every iteration (README.md:66, 88-98): the reference's
        # schedule; above the processed rate when settled states are skipped.
        "all_pairs_priced_gpairs_per_s": P * iters / (sat_ms / 1000.0) / 1e9,
        "roofline_frac": 9.0 * tested / iters / (sat_launch_ms / 1000.0) / 1e9 / HBM_PEAK_GBS,
        "identical_to_reference": _matches_reference(result, _golden(GOLDEN_OF.get(name, ""))),
        "_iterations_run": iterations_run, "_steps_tested": tested, "_iters": iters,
    }
    return out


def export_section(name):
    """The reference's /abstraction triplet for every pair copied to host memory
    (kernels.cu:806-824, 2869.32 ms for 401M pairs on its RTX 4090 run,
    README.md:72), with the digests checked against the reference build."""
    from gsc_amd.solver import Solver

    case = workload(name)
    with Solver(case, time_kernels=False) as solver:
        solver.build_abstraction()
        solver.solve()
        runs = []
        for _ in range(2):  # the first export also allocates the pinned staging
            t0 = time.perf_counter()
            res = solver.download(abstraction=True)
            runs.append(((time.perf_counter() - t0) * 1000.0, res.kernel_timings["result_memcpy_d2h_ms"]))
            last = res
            if len(runs) == 1:
                del res
    P = case.state_count() * case.input_count()
    fx = _golden(GOLDEN_OF.get(name, ""))
    same = None
    if fx is not None:
        same = all(_sha(getattr(last, a)) == fx["sha256"][k] for a, k in
                   (("abstraction_min_flat", "min_flat"), ("abstraction_max_flat", "max_flat"),
                    ("abstraction_valid", "pair_valid")))
    d2h = runs[-1][1]
    return {
        "workload": case.name, "pairs": P, "bytes": 9 * P,
        "download_ms_first": runs[0][1], "download_ms": d2h, "download_wall_ms": runs[-1][0],
        "GBps": 9 * P / (d2h / 1000.0) / 1e9,
        "reference_rtx4090_d2h_ms": 2869.32,
        "identical_to_reference": same,
        "how": "states' arrays plus the triplet, expanded on the device in 4M-pair slices, double-buffered through "
               "pinned staging into the caller's (fresh, pageable) arrays",
    }


def host_solver_section(name, threads=16):
    """The CLI's --compare mode on the same case: its multithreaded host solver
    (the --cpu mode, host/cpu_solver.cpp), then the gfx950 path, and the exact
    comparison of the five result arrays (main.cpp:189-211)."""
    import subprocess
    import tempfile

    cli = os.path.join(ROOT, "gpu-inner-outer_amd", "cuda_symbolic_control_cli")
    case = workload(name)
    threads = min(threads, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as tmp:
        cfg = os.path.join(tmp, "case.cfg")
        with open(cfg, "w") as f:
            f.write(case.to_cfg().replace("verbose = true", "verbose = false"))
        env = dict(os.environ, GSC_CPU_THREADS=str(threads))
        t0 = time.perf_counter()
        p = subprocess.run([cli, cfg, "--compare", "--output", os.path.join(tmp, "cmp.h5")], capture_output=True,
                           text=True, env=env, timeout=600)
        wall = time.perf_counter() - t0
    if p.returncode not in (0, 3):
        raise RuntimeError(f"--compare exited {p.returncode}: {p.stderr[-300:]}")
    # The [cpu] block's "abstraction" and "solve" lines (main.cpp:12-39).
    cpu_block = p.stdout.split("[cpu]", 1)[1].split("[gpu]", 1)[0]

    def field(label):
        for ln in cpu_block.splitlines():
            if ln.startswith(label):
                return float(ln.split(":", 1)[1].split()[0])
        return None

    ab, so = field("abstraction"), field("solve")
    return {
        "workload": case.name, "threads": threads, "abstraction_ms": ab, "solve_ms": so,
        "total_ms": None if ab is None or so is None else ab + so,
        "compare_wall_s": wall, "exact_match": "exact_match   : true" in p.stdout,
        "reference_cpu_ms": (_golden(GOLDEN_OF.get(name, "")) or {}).get("reference_cpu_ms"),
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
                    help="measure the dominant kernel's HBM bytes with rocprofv3 --pmc passes")
    ap.add_argument("--pmc-dir", default=os.path.join(ROOT, "gpurun_out", "bench_pmc"))
    ap.add_argument("--copy-calibration", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--batch-direct-only", action="store_true",
                    help="only the 32-case batch with directly launched stage kernels (for kernel profiles)")
    ap.add_argument("--no-batching", action="store_true", help="skip the 32-case batching comparison")
    ap.add_argument("--no-serving", action="store_true", help="skip the Poisson serving comparison")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the other schedules and workloads (sweep, published case, 128x128x64x16) and the export")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.batch_direct_only:
        print(json.dumps(batch_direct_section()))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1"}))
        sys.exit(2)

    print(f"[bench] workload {args.workload}", file=sys.stderr, flush=True)
    K = args.steps
    m = measure_workload(args.workload, K, args.warmup, sweep=args.sweep_all_pairs,
                         recompute=args.recompute_abstraction)
    tested, iters = m.pop("_steps_tested"), m.pop("_iters")
    iterations_run = m.pop("_iterations_run")
    sat_launch_ms = m["satisfaction_ms_per_iteration"]
    # Dominant kernel: the satisfaction pass. Algorithmic bytes are the
    # reference pair record (u32 min_flat + u32 max_flat + u8 valid = 9 B) per
    # pair the launch processes: every pair in the reference's schedule, the
    # pairs of still-unsettled states in the default one.
    units = tested / iters
    kname = "satisfy_recompute_kernel" if args.recompute_abstraction else "satisfy_kernel"
    roof = {"kernel": kname, "bound": "hbm", "launch_ms": sat_launch_ms,
            "algorithmic_bytes_per_launch": 9.0 * units, "pairs_processed_per_launch": units}
    roof["achieved"] = roof["algorithmic_bytes_per_launch"] / (sat_launch_ms / 1000.0) / 1e9
    roof.update({"peak": HBM_PEAK_GBS, "unit": "GB/s"})
    roof["frac"] = roof["achieved"] / roof["peak"]
    roof["traffic"] = None

    line = {
        "metric": METRIC,
        "value": m["processed_gpairs_per_s"],
        "unit": "Gpairs/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {k: m[k] for k in ("workload", "state_shape", "states", "inputs", "pairs", "iterations", "stop",
                                     "schedule", "abstraction")},
        "value_definition": "pairs tested by the satisfaction pass per second of its kernel time (HIP events)",
    }
    line["config"]["abstraction_dtype"] = "f64"
    for k in ("abstraction_ms", "abstraction_device_ms", "solve_ms", "total_ms", "solve_ms_with_kernel_events",
              "satisfaction_ms_per_iteration", "prefix_ms_per_iteration", "bookkeeping_ms_per_iteration",
              "pairs_tested_per_iteration", "boxes_counted_per_iteration", "test_lane_utilisation",
              "all_pairs_priced_gpairs_per_s", "identical_to_reference"):
        line[k] = m[k]
    line["iterations_run"] = iterations_run
    line["reference_published_rtx4090"] = {
        "workload": "hyperbolic_demo (70x70x63x13 x 100, README.md:44-102)", "abstraction_ms": 1002.53,
        "solve_ms": 214.936, "total_ms": 1217.47, "satisfaction_ms_per_iteration": 17.3, "gpairs_per_s": 23.2,
        "schedule": "sweep_all_pairs"}
    fx = _golden(GOLDEN_OF.get(args.workload, ""))
    if fx is not None and "reference_cpu_ms" in fx:
        r = fx["reference_cpu_ms"]
        line["reference_cpu_full_solve"] = {
            "abstraction_ms": r["abstraction"], "solve_ms": r["solve"], "total_ms": r["abstraction"] + r["solve"],
            "cores": 1, "where": "the build container (Intel Xeon @2.7 GHz), reference CPU build -O3",
            "source": "tests/golden (make_goldens.py --large)"}
    line["roofline"] = roof
    line["cpu_baseline"] = None
    if args.traffic == "live" and not args.recompute_abstraction:
        try:
            t = measure_traffic(args.workload, roof["kernel"], args.sweep_all_pairs, os.path.abspath(args.pmc_dir))
            roof["traffic"] = t["bytes_per_launch"]
            roof["traffic_detail"] = t
        except Exception as e:  # counters unavailable: traffic stays null, the reason is kept
            roof["traffic_error"] = str(e)[:300]
    if not args.no_extra:
        extra = {}
        for key, name, sweep, rec in (("sweep_all_pairs", args.workload, True, False),
                                      ("published_case", "hyperbolic_demo", False, False),
                                      ("published_case_sweep_all_pairs", "hyperbolic_demo", True, False),
                                      ("large_grid_stored", "unicycle_128x128x64x16", False, False),
                                      ("large_grid_recomputed", "unicycle_128x128x64x16", False, True)):
            print(f"[bench] section {key}", file=sys.stderr, flush=True)
            r = measure_workload(name, 3, 1, sweep=sweep, recompute=rec, log=False)
            for k in ("_steps_tested", "_iters", "_iterations_run", "ms_per_step", "solve_ms_with_kernel_events"):
                r.pop(k, None)
            extra[key] = r
        line["schedules_and_workloads"] = extra
        print("[bench] section export", file=sys.stderr, flush=True)
        line["abstraction_export"] = export_section(args.workload)
        print("[bench] section host solver", file=sys.stderr, flush=True)
        try:
            line["host_solver"] = host_solver_section(args.workload)
        except Exception as e:  # the CLI or its HDF5 library unavailable: the reason is kept
            line["host_solver"] = {"error": str(e)[:300]}
    if not args.no_batching:
        line["batching"] = batching_section()
    if not args.no_serving:
        line["serving"] = serving_section()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(workload(args.workload), args.cpu_budget)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
