"""The CLI's --cpu and --compare modes (src/main.cpp:157-211) on the host solver
of host/cpu_solver.cpp: its results against the oracle on seeded cases and
against the reference build's goldens at full size, the reference's CPU file
layout (src/result_io.cpp:238-362) and printout, and --compare's exit codes."""
import hashlib
import json
import os
import subprocess

import numpy as np
import pytest

import casegen
from conftest import GOLDEN
from gsc_amd import cases
from test_cli import CLI, H5DUMP, LAYOUT, _cli_or_skip, _h5_array, _h5_layout

KEYS = (("valid_mask", np.uint8), ("candidate_mask", np.uint8), ("reachable_mask", np.uint8),
        ("controller", np.uint32), ("reach_step", np.uint32))


def _run(case_text, tmp_path, *mode, env=None, name="c"):
    _cli_or_skip()
    cfg = tmp_path / f"{name}.cfg"
    cfg.write_text(case_text)
    out = tmp_path / f"{name}.h5"
    e = dict(os.environ)
    if env:
        e.update(env)
    p = subprocess.run([CLI, str(cfg), *mode, "--output", str(out)], capture_output=True, text=True, timeout=900,
                       env=e)
    return p, out


def _arrays(out, label, tmp_path):
    if H5DUMP is None:
        pytest.skip("h5dump not available to read the result file")
    return {k: _h5_array(str(out), f"/result/{label}/{k}", t, str(tmp_path)) for k, t in KEYS}


def _assert_same(got, ref, label):
    for k, _ in KEYS:
        assert np.array_equal(got[k], ref[k]), f"{label}: {k} differs at {int((got[k] != ref[k]).sum())} states"


def _iteration_lines(stdout):
    return [line for line in stdout.splitlines() if line.startswith("iter ")]


@pytest.mark.parametrize("seed", range(40))
def test_cpu_mode_matches_oracle_on_random_cases(oracle, tmp_path, seed):
    """Every wrap axis, both constraints, wrapped and full-range θ boxes, boxes
    leaving the domain, odd substep counts and the extension keys."""
    case = casegen.small_random_case(seed)
    p, out = _run(case.to_cfg(), tmp_path, "--cpu")
    try:
        oracle.masks(case)
    except ValueError as e:
        assert p.returncode == 1 and str(e) in p.stderr, p.stderr
        return
    assert p.returncode == 0, p.stderr
    ref = oracle.run(case, keep_abstraction=False)
    _assert_same(_arrays(out, "cpu", tmp_path), ref, case.name)
    assert f"iterations    : {ref['iterations']}" in p.stdout
    assert f"message       : {ref['message']}" in p.stdout


@pytest.mark.parametrize("which", ["default", "wraparound", "smoke", "hyperbolic"])
def test_cpu_mode_canned_cases(oracle, tmp_path, which):
    case = cases.builtin(which)
    p, out = _run(case.to_cfg(), tmp_path, "--cpu")
    assert p.returncode == 0, p.stderr
    got = _arrays(out, "cpu", tmp_path)
    gold = os.path.join(GOLDEN, f"builtin_{which}.npz")
    if os.path.exists(gold):
        ref = dict(np.load(gold))
    else:
        ref = oracle.run(case, keep_abstraction=False)
    _assert_same(got, ref, which)


def test_cpu_mode_overflow_and_wrap_x(oracle, tmp_path):
    """x is the wrap axis and most boxes wrap around it."""
    case = casegen.overflow_regrowth_case()
    p, out = _run(case.to_cfg(), tmp_path, "--cpu")
    assert p.returncode == 0, p.stderr
    _assert_same(_arrays(out, "cpu", tmp_path), oracle.run(case, keep_abstraction=False), case.name)


@pytest.mark.parametrize("golden", ["unicycle_30x30x24x10_T0.5", "unicycle_30x30x24x10_T0.1", "hyperbolic_demo",
                                    "unicycle_91x105x42x10"])
def test_cpu_mode_matches_reference_goldens(tmp_path, golden):
    """Digests and per-iteration counts written by the reference's own CPU
    build; the last two are the 401.31M-pair cases."""
    g = json.load(open(os.path.join(GOLDEN, golden + ".json")))
    # verbose on, for the per-iteration lines (it changes nothing else).
    p, out = _run(g["cfg"].replace("verbose = false", "verbose = true"), tmp_path, "--cpu")
    assert p.returncode == 0, p.stderr
    got = _arrays(out, "cpu", tmp_path)
    for k, _ in KEYS:
        assert hashlib.sha256(np.ascontiguousarray(got[k]).tobytes()).hexdigest() == g["sha256"][k], k
    assert f"iterations    : {g['iterations']}" in p.stdout
    assert f"message       : {g['message']}" in p.stdout
    assert f"reachable     : {g['reachable_states']}" in p.stdout
    # reachability.hpp:113-128: one line per iteration.
    lines = _iteration_lines(p.stdout)
    assert [int(line.split("new_reachable=")[1].split()[0]) for line in lines] == g["newly_reachable"]


def test_cpu_mode_thread_count_does_not_change_results(tmp_path):
    case = cases.unicycle_30x30x24x10(sample_time=0.5)
    p1, out1 = _run(case.to_cfg(), tmp_path, "--cpu", env={"GSC_CPU_THREADS": "1"}, name="t1")
    p7, out7 = _run(case.to_cfg(), tmp_path, "--cpu", env={"GSC_CPU_THREADS": "7"}, name="t7")
    assert p1.returncode == 0 and p7.returncode == 0, p1.stderr + p7.stderr
    _assert_same(_arrays(out1, "cpu", tmp_path), _arrays(out7, "cpu", tmp_path), "1 vs 7 threads")
    assert _iteration_lines(p1.stdout) == _iteration_lines(p7.stdout)


def test_cpu_mode_is_the_default_and_prints_the_reference_lines(tmp_path):
    case = cases.builtin("default")
    _cli_or_skip()
    cfg = tmp_path / "d.cfg"
    cfg.write_text(case.to_cfg())
    p = subprocess.run([CLI, str(cfg)], capture_output=True, text=True, cwd=tmp_path, timeout=300)
    assert p.returncode == 0, p.stderr
    S, I = case.state_count(), case.input_count()
    # abstraction.hpp:291-304 and main.cpp:12-39, 157-162
    for line in (f"CPU: Building valid mask for {S} states...", f"CPU: Building candidate mask for {S} states...",
                 f"CPU: Starting abstraction phase for {S} states and {I} inputs...",
                 f"CPU: Valid pairs: {S * I}", "[cpu]", f"states        : {S}", f"pairs         : {S * I}",
                 "result file   : results/unicycle_default_cpu.h5"):
        assert line in p.stdout, line
    assert (tmp_path / "results" / "unicycle_default_cpu.h5").exists()


def test_cpu_file_layout(tmp_path):
    if H5DUMP is None:
        pytest.skip("h5dump not available")
    case = cases.builtin("wraparound")
    p, out = _run(case.to_cfg(), tmp_path, "--cpu")
    assert p.returncode == 0, p.stderr
    lay = _h5_layout(str(out))
    S = case.state_count()
    expect = {k.replace("/gpu/", "/cpu/"): v for k, v in LAYOUT.items()
              if k.startswith(("/meta/", "/state/", "/input/", "/config/", "/summary/gpu/", "/result/gpu/"))
              and k != "/summary/gpu/pair_count"}
    sizes = {"S": S, "I": case.input_count(), "O": len(case.obstacles)}
    for dset, (tword, shape) in expect.items():
        assert dset in lay, f"missing dataset {dset}"
        assert lay[dset]["type"].startswith(tword), dset
        want = None if shape is None else tuple(sizes[d] if isinstance(d, str) else d for d in shape)
        assert lay[dset]["shape"] == want, dset
    # No GPU groups in a CPU file (result_io.cpp:238-316).
    assert not any(k.startswith(("/abstraction/", "/iterations/", "/timings/", "/summary/gpu", "/result/gpu"))
                   for k in lay)


def test_cpu_mode_config_errors(tmp_path):
    p, _ = _run("name = e\nstate_eta = 0.5, 0, 0.5, 0.5\n", tmp_path, "--cpu")
    assert p.returncode == 1 and "state eta must be positive" in p.stderr
    p, _ = _run("name = e\ninput_lb = 1, -1\ninput_ub = 1, 1\n", tmp_path, "--cpu")
    assert p.returncode == 1 and "input bounds must satisfy ub > lb" in p.stderr


def test_compare_without_a_device_writes_the_cpu_run_and_exits_2(tmp_path):
    """main.cpp:194-200: no GPU run, the CPU result is still written."""
    from gsc_amd import capi

    try:
        has_device = capi.load().gsc_device_count() > 0
    except Exception:
        has_device = False
    if has_device:
        pytest.skip("a device is present; the GPU test covers --compare")
    case = cases.builtin("default")
    p, out = _run(case.to_cfg(), tmp_path, "--compare")
    assert p.returncode == 2, p.stdout + p.stderr
    assert "gpu compare unavailable" in p.stderr
    assert out.exists() and "[cpu]" in p.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("which", ["default", "wraparound", "smoke", "hyperbolic", "c1", "seed:3", "seed:14"])
def test_compare_mode_exact_match(gpu, tmp_path, which):
    """--compare runs the host solver and the gfx950 path and exits 0 only when
    the five arrays agree (main.cpp:189-211)."""
    if which.startswith("seed:"):
        case = casegen.small_random_case(int(which.split(":")[1]))
    elif which == "c1":
        case = cases.unicycle_30x30x24x10(sample_time=0.5)
    else:
        case = cases.builtin(which)
    print(f"[case] CLI --compare {which}", flush=True)
    p, out = _run(case.to_cfg(), tmp_path, "--compare")
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr
    assert "exact_match   : true" in p.stdout
    assert "[cpu]" in p.stdout and "[gpu]" in p.stdout and "[GPU Detailed Timing]" in p.stdout
    got_cpu, got_gpu = _arrays(out, "cpu", tmp_path), _arrays(out, "gpu", tmp_path)
    _assert_same(got_gpu, got_cpu, which)
