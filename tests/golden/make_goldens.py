"""Writes the golden fixtures in tests/golden/ from the reference's own CPU build.

Run in the build container (needs /root/reference and `make -C oracle ref`):
    python tests/golden/make_goldens.py [--large]

For every case, oracle/_ref/ref_harness (load_case_config -> prepare_case_cpu
-> solve_backward_cpu, compiled from the reference sources) is run and the
fixture records SHA-256 digests of every output array, the per-iteration
counts and the stop reason. Tiny cases also keep their full arrays (.npz).
`--large` adds the 401,310,000-pair cases (several minutes each).
`--batch` and `--serving` write digests and per-iteration counts for every
case of the full-size batching workload (32 cases of 5e4-1e6 cells) and the
serving workload (48 requests of 2e4-4e6 cells), several reference processes
at a time (`--jobs`); with `--only NAME` the small fixtures are left as they are.

The published case is additionally tied to the two artefacts the reference
ships: the README run (README.md:44-98: counts per iteration) and the
plots/reachable_set.mat digest of its reachable set (reachable_mask_3d, a
70x70x63 projection over the v axis).
"""
import argparse
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "gpu-inner-outer_amd", "python"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import gsc_oracle  # noqa: E402
from casegen import small_random_case  # noqa: E402
from gsc_amd import cases  # noqa: E402

ARRAYS = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step", "min_flat", "max_flat",
          "pair_valid")


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def fixture(case, ref, keep_arrays=False, key=None):
    key = key or case.name
    out = {
        "name": case.name,
        "cfg": case.to_cfg(),
        "iterations": ref["iterations"],
        "converged": ref["converged"],
        "message": ref["message"],
        "reachable_states": ref["reachable_states"],
        "certified_candidate_states": ref["certified_candidate_states"],
        "newly_reachable": [int(v) for v in ref["newly_reachable"]],
        "newly_certified": [int(v) for v in ref["newly_certified"]],
        "valid_states": int(ref["valid_mask"].sum()),
        "candidate_states": int(ref["candidate_mask"].sum()),
        "valid_pairs": int(ref["pair_valid"].sum()),
        "sha256": {k: sha(ref[k]) for k in ARRAYS},
        "source": "reference CPU build (oracle/_ref/ref_harness)",
    }
    if keep_arrays:
        np.savez_compressed(os.path.join(HERE, f"{key}.npz"), **{k: ref[k] for k in ARRAYS})
    return out


def write(name, obj):
    with open(os.path.join(HERE, f"{name}.json"), "w") as f:
        json.dump(obj, f, indent=1, sort_keys=True)
    print("wrote", name, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--large", action="store_true")
    ap.add_argument("--only", default=None)
    ap.add_argument("--batch", action="store_true", help="the 32 full-size batching cases")
    ap.add_argument("--serving", action="store_true", help="the 48 full-size serving requests")
    ap.add_argument("--jobs", type=int, default=6, help="reference processes at a time for --batch / --serving")
    args = ap.parse_args()
    if not gsc_oracle.ref_available():
        raise SystemExit("build the reference harness first: make -C oracle ref")

    # The canned cases come from the reference's own make_*_case() (not from our
    # transcription of them), so these fixtures also pin gsc_config_builtin.
    for which, max_it in (("default", None), ("wraparound", None), ("smoke", None), ("hyperbolic", 3)):
        key = f"builtin_{which}"
        if args.only and args.only != key:
            continue
        case = cases.builtin(which)
        ref = gsc_oracle.run_ref(case, builtin=which, max_iterations=max_it)
        fx = fixture(case, ref, keep_arrays=case.state_count() * case.input_count() <= 20000, key=key)
        fx["max_iterations_override"] = max_it
        write(key, fx)

    small = {
        "unicycle_30x30x24x10_T0.1": cases.unicycle_30x30x24x10(sample_time=0.1),
        "unicycle_30x30x24x10_T0.5": cases.unicycle_30x30x24x10(sample_time=0.5),
    }
    for key, case in small.items():
        if args.only and args.only != key:
            continue
        ref = gsc_oracle.run_ref(case)
        write(key, fixture(case, ref, keep_arrays=case.state_count() * case.input_count() <= 20000, key=key))

    if not args.only or args.only == "small_random":
        rnd = {}
        for seed in range(24):
            case = small_random_case(seed)
            try:
                ref = gsc_oracle.run_ref(case)
            except RuntimeError as e:
                rnd[str(seed)] = {"name": case.name, "error": str(e).replace("error: ", "")}
                continue
            rnd[str(seed)] = fixture(case, ref)
        write("small_random", rnd)

    if args.large and (not args.only or args.only == "hyperbolic_demo"):
        case = cases.hyperbolic_demo()
        ref = gsc_oracle.run_ref(case, timeout=7200)
        fx = fixture(case, ref)
        fx["reference_cpu_ms"] = {"abstraction": ref["abstraction_ms"], "solve": ref["solve_ms"]}
        # The reference's committed fixture: plots/reachable_set.mat, reachable_mask_3d.
        mat = "/root/reference/plots/reachable_set.mat"
        if os.path.exists(mat):
            import scipy.io

            m = scipy.io.loadmat(mat)["reachable_mask_3d"].astype(np.uint8)
            fx["mat_reachable_mask_3d"] = {"shape": list(m.shape), "sum": int(m.sum()),
                                           "sha256_C_order": sha(np.ascontiguousarray(m))}
        write("hyperbolic_demo", fx)

    if args.large and (not args.only or args.only == "unicycle_91x105x42x10"):
        case = cases.unicycle_91x105x42x10()
        ref = gsc_oracle.run_ref(case, timeout=7200)
        fx = fixture(case, ref)
        fx["reference_cpu_ms"] = {"abstraction": ref["abstraction_ms"], "solve": ref["solve_ms"]}
        write("unicycle_91x105x42x10", fx)

    # Full-size batching and serving workloads: every case of
    # batch_instances(32, seed=3) (5e4-1e6 cells) and of serving_requests(48,
    # seed=5) (2e4-4e6 cells), each solved by the reference build on its own.
    if args.batch and (not args.only or args.only == "batch_instances_32_s3"):
        write("batch_instances_32_s3", many_fixtures(cases.batch_instances(32, seed=3), args.jobs))
    if args.serving and (not args.only or args.only == "serving_requests_48_s5"):
        write("serving_requests_48_s5", many_fixtures(cases.serving_requests(48, seed=5), args.jobs))


def _one_fixture(case):
    ref = gsc_oracle.run_ref(case, timeout=4 * 3600)
    fx = fixture(case, ref)
    fx["reference_cpu_ms"] = {"abstraction": ref["abstraction_ms"], "solve": ref["solve_ms"]}
    print(f"  {case.name}: {ref['iterations']} iterations, {ref['message']}", flush=True)
    return case.name, fx


def many_fixtures(case_list, jobs):
    """Reference solves of many cases, one single-threaded reference process per
    case, `jobs` at a time; largest first so the long ones start early."""
    from concurrent.futures import ThreadPoolExecutor

    order = sorted(case_list, key=lambda c: -c.state_count() * c.input_count())
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        done = dict(ex.map(_one_fixture, order))
    return {c.name: done[c.name] for c in case_list}


if __name__ == "__main__":
    main()
