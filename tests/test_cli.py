"""The drop-in CLI `cuda_symbolic_control_cli <cfg> --gpu` (src/main.cpp:125-218)
and its HDF5 file (src/result_io.cpp:364-517)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gsc_amd import cases

CLI = os.path.join(ROOT, "gpu-inner-outer_amd", "cuda_symbolic_control_cli")
H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)

# Every dataset of a --gpu result file: path -> (HDF5 type class word, shape or None for scalar).
LAYOUT = {
    "/meta/case_name": ("STRING", None), "/meta/invalid_input": ("32-bit", None),
    "/meta/unreachable_step": ("32-bit", None), "/meta/run_count": ("32-bit", None),
    "/state/lb": ("64-bit", (4,)), "/state/ub": ("64-bit", (4,)), "/state/eta": ("64-bit", (4,)),
    "/state/shape": ("32-bit", (4,)), "/state/strides": ("64-bit", (4,)), "/state/total_size": ("64-bit", None),
    "/state/wrap_dim": ("32-bit", None),
    "/input/lb": ("64-bit", (2,)), "/input/ub": ("64-bit", (2,)), "/input/eta": ("64-bit", (2,)),
    "/input/shape": ("32-bit", (2,)), "/input/strides": ("64-bit", (2,)), "/input/total_size": ("64-bit", None),
    "/input/centers": ("64-bit", ("I", 2)),
    "/config/candidate_lb": ("64-bit", (4,)), "/config/candidate_ub": ("64-bit", (4,)),
    "/config/map_lb": ("64-bit", (4,)), "/config/map_ub": ("64-bit", (4,)),
    "/config/disturbance_half_width": ("64-bit", (4,)), "/config/sample_time": ("64-bit", None),
    "/config/integration_substeps": ("32-bit", None), "/config/max_iterations": ("32-bit", None),
    "/config/verbose": ("8-bit", None), "/config/obstacles": ("64-bit", ("O", 8)),
    "/summary/gpu/iterations": ("32-bit", None), "/summary/gpu/converged": ("8-bit", None),
    "/summary/gpu/message": ("STRING", None), "/summary/gpu/reachable_states": ("64-bit", None),
    "/summary/gpu/certified_candidate_states": ("64-bit", None), "/summary/gpu/abstraction_ms": ("64-bit", None),
    "/summary/gpu/solve_ms": ("64-bit", None), "/summary/gpu/pair_count": ("64-bit", None),
    "/result/gpu/valid_mask": ("8-bit", ("S",)), "/result/gpu/candidate_mask": ("8-bit", ("S",)),
    "/result/gpu/reachable_mask": ("8-bit", ("S",)), "/result/gpu/controller": ("32-bit", ("S",)),
    "/result/gpu/reach_step": ("32-bit", ("S",)),
    "/abstraction/pair_count": ("64-bit", None), "/abstraction/min_flat": ("32-bit", ("P",)),
    "/abstraction/max_flat": ("32-bit", ("P",)), "/abstraction/valid": ("8-bit", ("P",)),
    "/iterations/iteration": ("32-bit", ("N",)), "/iterations/newly_reachable": ("64-bit", ("N",)),
    "/iterations/newly_certified": ("64-bit", ("N",)), "/iterations/total_reachable": ("64-bit", ("N",)),
    "/iterations/total_certified": ("64-bit", ("N",)), "/iterations/iteration_ms": ("64-bit", ("N",)),
    "/iterations/prefix_build_ms": ("64-bit", ("N",)), "/iterations/satisfaction_check_ms": ("64-bit", ("N",)),
    "/iterations/reduction_ms": ("64-bit", ("N",)),
}
TIMINGS = ["abstraction_kernel_ms", "abstraction_memcpy_h2d_ms", "scatter_mask_ms", "scan_axis0_ms", "scan_axis1_ms",
           "scan_axis2_ms", "scan_axis3_ms", "prefix_build_total_ms", "pair_satisfaction_ms", "reduce_inputs_ms",
           "iteration_memcpy_ms", "result_memcpy_d2h_ms", "total_kernel_ms", "total_memcpy_ms", "total_compute_ms"]
for _t in TIMINGS:
    LAYOUT[f"/timings/{_t}"] = ("64-bit", None)


def _cli_or_skip():
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} not built (make -C gpu-inner-outer_amd cli)")


def test_cli_usage_and_errors(tmp_path):
    _cli_or_skip()
    p = subprocess.run([CLI], capture_output=True, text=True)
    assert p.returncode == 1 and "usage" in p.stderr
    p = subprocess.run([CLI, str(tmp_path / "missing.cfg"), "--gpu"], capture_output=True, text=True)
    assert p.returncode == 1 and "failed to open config" in p.stderr
    bad = tmp_path / "bad.cfg"
    bad.write_text("colour = blue\n")
    p = subprocess.run([CLI, str(bad), "--gpu"], capture_output=True, text=True)
    assert p.returncode == 1 and "unknown config key: colour" in p.stderr
    p = subprocess.run([CLI, str(bad), "--gpu", "--output"], capture_output=True, text=True)
    assert p.returncode == 1 and "--output requires a file path" in p.stderr
    ok = tmp_path / "ok.cfg"
    ok.write_text("name = x\n")
    p = subprocess.run([CLI, str(ok), "--frobnicate"], capture_output=True, text=True)
    assert p.returncode == 1 and "unknown argument: --frobnicate" in p.stderr


def _h5_layout(path):
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    found = {}
    stack = []
    for line in out.splitlines():
        m = re.match(r'\s*GROUP "([^"]+)"', line)
        if m:
            stack.append(m.group(1))
            continue
        m = re.match(r'\s*DATASET "([^"]+)"', line)
        if m:
            path_parts = [g for g in stack if g != "/"] + [m.group(1)]
            cur = "/" + "/".join(path_parts)
            found[cur] = {"type": None, "shape": None}
            continue
        if re.match(r"\s*}\s*$", line) and stack and not found.get("__in_ds__"):
            pass
        m = re.match(r"\s*DATATYPE\s+(.*)", line)
        if m and found:
            last = list(found)[-1]
            found[last]["type"] = m.group(1)
        m = re.match(r"\s*DATASPACE\s+(SCALAR|SIMPLE\s*\{\s*\(([^)]*)\))", line)
        if m and found:
            last = list(found)[-1]
            found[last]["shape"] = None if m.group(1) == "SCALAR" else tuple(
                int(v) for v in m.group(2).split(",") if v.strip())
    return found


def _h5_array(path, dset, dtype, tmp):
    raw = os.path.join(tmp, dset.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", raw, path], capture_output=True, check=True)
    return np.fromfile(raw, dtype=dtype)


@pytest.mark.gpu
def test_cli_gpu_run_writes_reference_layout(gpu, oracle, tmp_path):
    _cli_or_skip()
    case = cases.unicycle_30x30x24x10(sample_time=0.5)
    cfg = tmp_path / "c1.cfg"
    cfg.write_text(case.reference_cfg())
    out = tmp_path / "res" / "c1.h5"
    print(f"[case] CLI {case.name}", flush=True)
    p = subprocess.run([CLI, str(cfg), "--gpu", "--output", str(out)], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr
    ref = oracle.run(case)
    S, I = case.state_count(), case.input_count()
    assert f"states        : {S}" in p.stdout
    assert f"pairs         : {S * I}" in p.stdout
    assert f"reachable     : {int(ref['reachable_mask'].sum())}" in p.stdout
    assert f"iterations    : {ref['iterations']}" in p.stdout
    assert f"message       : {ref['message']}" in p.stdout
    assert "=== Per-Iteration Breakdown (first 5 and last 5) ===" in p.stdout
    assert os.path.exists(out)
    if H5DUMP is None:
        pytest.skip("h5dump not available to inspect the file")
    lay = _h5_layout(str(out))
    N = ref["iterations"]
    sizes = {"S": S, "P": S * I, "I": I, "N": N, "O": len(case.obstacles)}
    for dset, (tword, shape) in LAYOUT.items():
        assert dset in lay, f"missing dataset {dset}"
        assert tword in lay[dset]["type"], f"{dset}: type {lay[dset]['type']}"
        want = None if shape is None else tuple(sizes[d] if isinstance(d, str) else d for d in shape)
        assert lay[dset]["shape"] == want, f"{dset}: shape {lay[dset]['shape']} != {want}"
    tmp = str(tmp_path)
    for k, t in (("valid_mask", np.uint8), ("candidate_mask", np.uint8), ("reachable_mask", np.uint8),
                 ("controller", np.uint32), ("reach_step", np.uint32)):
        assert np.array_equal(_h5_array(str(out), f"/result/gpu/{k}", t, tmp), ref[k]), k
    for k, rk, t in (("min_flat", "min_flat", np.uint32), ("max_flat", "max_flat", np.uint32),
                     ("valid", "pair_valid", np.uint8)):
        assert np.array_equal(_h5_array(str(out), f"/abstraction/{k}", t, tmp), ref[rk]), k
    nr = _h5_array(str(out), "/iterations/newly_reachable", np.uint64, tmp)
    assert [int(v) for v in nr] == [int(v) for v in ref["newly_reachable"]]


@pytest.mark.gpu
def test_cli_default_result_path(gpu, tmp_path):
    _cli_or_skip()
    cfg = tmp_path / "n.cfg"
    cfg.write_text("name = my case/1\nmax_iterations = 2\n")
    p = subprocess.run([CLI, str(cfg), "--gpu", "--no-abstraction-export"], capture_output=True, text=True,
                       cwd=tmp_path, timeout=120)
    assert p.returncode == 0, p.stderr
    assert (tmp_path / "results" / "my_case_1_gpu.h5").exists()
