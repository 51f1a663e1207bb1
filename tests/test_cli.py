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

# Every dataset of a --gpu result file: path -> (HDF5 type, shape or None for
# scalars); the types are those of the reference's C++ values (u32 -> U32LE,
# double -> F64LE, int -> I32LE, ...).
U8, U32, U64, I32, F64, STR = "H5T_STD_U8LE", "H5T_STD_U32LE", "H5T_STD_U64LE", "H5T_STD_I32LE", "H5T_IEEE_F64LE", \
    "H5T_STRING"
LAYOUT = {
    "/meta/case_name": (STR, None), "/meta/invalid_input": (U32, None),
    "/meta/unreachable_step": (U32, None), "/meta/run_count": (U32, None),
    "/state/lb": (F64, (4,)), "/state/ub": (F64, (4,)), "/state/eta": (F64, (4,)),
    "/state/shape": (U32, (4,)), "/state/strides": (U64, (4,)), "/state/total_size": (U64, None),
    "/state/wrap_dim": (I32, None),
    "/input/lb": (F64, (2,)), "/input/ub": (F64, (2,)), "/input/eta": (F64, (2,)),
    "/input/shape": (U32, (2,)), "/input/strides": (U64, (2,)), "/input/total_size": (U64, None),
    "/input/centers": (F64, ("I", 2)),
    "/config/candidate_lb": (F64, (4,)), "/config/candidate_ub": (F64, (4,)),
    "/config/map_lb": (F64, (4,)), "/config/map_ub": (F64, (4,)),
    "/config/disturbance_half_width": (F64, (4,)), "/config/sample_time": (F64, None),
    "/config/integration_substeps": (U32, None), "/config/max_iterations": (U32, None),
    "/config/verbose": (U8, None), "/config/obstacles": (F64, ("O", 8)),
    "/summary/gpu/iterations": (U32, None), "/summary/gpu/converged": (U8, None),
    "/summary/gpu/message": (STR, None), "/summary/gpu/reachable_states": (U64, None),
    "/summary/gpu/certified_candidate_states": (U64, None), "/summary/gpu/abstraction_ms": (F64, None),
    "/summary/gpu/solve_ms": (F64, None), "/summary/gpu/pair_count": (U64, None),
    "/result/gpu/valid_mask": (U8, ("S",)), "/result/gpu/candidate_mask": (U8, ("S",)),
    "/result/gpu/reachable_mask": (U8, ("S",)), "/result/gpu/controller": (U32, ("S",)),
    "/result/gpu/reach_step": (U32, ("S",)),
    "/abstraction/pair_count": (U64, None), "/abstraction/min_flat": (U32, ("P",)),
    "/abstraction/max_flat": (U32, ("P",)), "/abstraction/valid": (U8, ("P",)),
    "/iterations/iteration": (U32, ("N",)), "/iterations/newly_reachable": (U64, ("N",)),
    "/iterations/newly_certified": (U64, ("N",)), "/iterations/total_reachable": (U64, ("N",)),
    "/iterations/total_certified": (U64, ("N",)), "/iterations/iteration_ms": (F64, ("N",)),
    "/iterations/prefix_build_ms": (F64, ("N",)), "/iterations/satisfaction_check_ms": (F64, ("N",)),
    "/iterations/reduction_ms": (F64, ("N",)),
}
TIMINGS = ["abstraction_kernel_ms", "abstraction_memcpy_h2d_ms", "scatter_mask_ms", "scan_axis0_ms", "scan_axis1_ms",
           "scan_axis2_ms", "scan_axis3_ms", "prefix_build_total_ms", "pair_satisfaction_ms", "reduce_inputs_ms",
           "iteration_memcpy_ms", "result_memcpy_d2h_ms", "total_kernel_ms", "total_memcpy_ms", "total_compute_ms"]
for _t in TIMINGS:
    LAYOUT[f"/timings/{_t}"] = (F64, None)


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
    """{dataset path: {"type": DATATYPE text, "shape": dims or None}} from `h5dump -H`.
    h5dump indents each nesting level by the same step, so a dataset's groups
    are the GROUP lines above it at smaller indents."""
    out = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, check=True).stdout
    found = {}
    groups = {}  # indent -> group name
    cur = None
    for line in out.splitlines():
        indent = len(line) - len(line.lstrip(" "))
        m = re.match(r'\s*GROUP "([^"]+)"', line)
        if m:
            groups = {k: v for k, v in groups.items() if k < indent}
            groups[indent] = m.group(1)
            cur = None
            continue
        m = re.match(r'\s*DATASET "([^"]+)"', line)
        if m:
            parents = [groups[k] for k in sorted(groups) if k < indent and groups[k] != "/"]
            cur = "/" + "/".join(parents + [m.group(1)])
            found[cur] = {"type": None, "shape": None}
            continue
        if cur is None:
            continue
        m = re.match(r"\s*DATATYPE\s+(.*)", line)
        if m:
            found[cur]["type"] = m.group(1)
        m = re.match(r"\s*DATASPACE\s+(SCALAR|SIMPLE\s*\{\s*\(([^)]*)\))", line)
        if m:
            found[cur]["shape"] = None if m.group(1) == "SCALAR" else tuple(
                int(v) for v in m.group(2).split(",") if v.strip())
    return found


def check_layout(path, sizes):
    lay = _h5_layout(path)
    for dset, (tword, shape) in LAYOUT.items():
        assert dset in lay, f"missing dataset {dset}"
        assert lay[dset]["type"] is not None and lay[dset]["type"].startswith(tword), f"{dset}: type {lay[dset]['type']}"
        want = None if shape is None else tuple(sizes[d] if isinstance(d, str) else d for d in shape)
        assert lay[dset]["shape"] == want, f"{dset}: shape {lay[dset]['shape']} != {want}"
    return lay


def _h5_array(path, dset, dtype, tmp):
    raw = os.path.join(tmp, dset.strip("/").replace("/", "_") + ".bin")
    subprocess.run([H5DUMP, "-d", dset, "-b", "LE", "-o", raw, path], capture_output=True, check=True)
    return np.fromfile(raw, dtype=dtype)


def _h5_text(path, dset):
    out = subprocess.run([H5DUMP, "-d", dset, path], capture_output=True, text=True, check=True).stdout
    m = re.search(r"DATA\s*\{\s*(.*?)\s*\}", out, re.S)
    return m.group(1)


def test_hdf5_writer_layout_and_values(tmp_path):
    """The writer alone, fed a hand-made report (tests/cpp/h5_writer_fixture.cpp)."""
    if H5DUMP is None or not os.path.exists("/opt/conda/include/hdf5.h"):
        pytest.skip("HDF5 headers or h5dump not available")
    exe = tmp_path / "h5_writer_fixture"
    host = os.path.join(ROOT, "gpu-inner-outer_amd", "host")
    subprocess.run(["g++", "-O1", "-std=c++17", "-DGSC_HAS_HDF5=1", "-I/opt/conda/include", f"-I{host}",
                    os.path.join(ROOT, "tests", "cpp", "h5_writer_fixture.cpp"), os.path.join(host, "result_io.cpp"),
                    "/opt/conda/lib/libhdf5.so", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu:/opt/conda/lib",
                    "-o", str(exe)], check=True, capture_output=True)
    out = tmp_path / "sub" / "f.h5"
    subprocess.run([str(exe), str(out)], check=True)
    S, P = 48, 288
    check_layout(str(out), {"S": S, "P": P, "I": 6, "N": 3, "O": 2})
    t = str(tmp_path)
    s = np.arange(S)
    assert np.array_equal(_h5_array(str(out), "/result/gpu/valid_mask", np.uint8, t), (s % 5 != 0).astype(np.uint8))
    assert np.array_equal(_h5_array(str(out), "/result/gpu/candidate_mask", np.uint8, t), (s % 7 == 1).astype(np.uint8))
    assert np.array_equal(_h5_array(str(out), "/result/gpu/reachable_mask", np.uint8, t), (s % 3 == 0).astype(np.uint8))
    assert np.array_equal(_h5_array(str(out), "/result/gpu/controller", np.uint32, t),
                          (s.astype(np.uint64) * 2654435761 % 2**32).astype(np.uint32))
    assert np.array_equal(_h5_array(str(out), "/result/gpu/reach_step", np.uint32, t), (s % 4).astype(np.uint32))
    p = np.arange(P)
    assert np.array_equal(_h5_array(str(out), "/abstraction/min_flat", np.uint32, t), (p % S).astype(np.uint32))
    assert np.array_equal(_h5_array(str(out), "/abstraction/max_flat", np.uint32, t), (p * 7 % S).astype(np.uint32))
    assert np.array_equal(_h5_array(str(out), "/abstraction/valid", np.uint8, t), (p % 4 != 0).astype(np.uint8))
    assert list(_h5_array(str(out), "/iterations/newly_reachable", np.uint64, t)) == [10, 20, 30]
    assert list(_h5_array(str(out), "/iterations/iteration", np.uint32, t)) == [1, 2, 3]
    assert list(_h5_array(str(out), "/state/shape", np.uint32, t)) == [4, 3, 2, 2]
    assert list(_h5_array(str(out), "/state/strides", np.uint64, t)) == [1, 4, 12, 24]
    # one row per obstacle: lb[0..3] then ub[0..3] (result_io.cpp:228-236)
    assert list(_h5_array(str(out), "/config/obstacles", np.float64, t)) == [0, 1, 2, 3, 0.5, 1.5, 2.5, 3.5] * 2
    assert '"fixture case"' in _h5_text(str(out), "/meta/case_name")
    assert '"all candidate states obtained a controller"' in _h5_text(str(out), "/summary/gpu/message")
    assert _h5_text(str(out), "/meta/run_count").split(":")[-1].strip() == "1"
    # input centres: lb + (idx + 0.5) * eta, x fastest (grid.hpp:68-78)
    c = _h5_array(str(out), "/input/centers", np.float64, t).reshape(6, 2)
    lb, eta = (-0.5, -1.0), (0.5, 1.0)
    want = [[lb[0] + (k % 2 + 0.5) * eta[0], lb[1] + (k // 2 + 0.5) * eta[1]] for k in range(6)]
    assert np.array_equal(c, np.array(want))


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
    N = ref["iterations"]
    check_layout(str(out), {"S": S, "P": S * I, "I": I, "N": N, "O": len(case.obstacles)})
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
