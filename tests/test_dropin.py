"""The binary drop-in: the reference's own programs linked against libgsc_hip.so.

The reference selects its GPU backend at link time (CMakeLists.txt:11-40): a
static CUDA library exporting gsc::cuda_backend_compiled() and
gsc::run_case_cuda_u32(const CaseConfig&) (cuda_api.hpp:87-88). This backend
ships libgsc_hip.so with the same two symbols and the same struct layouts.
`make -C oracle dropin` compiles the reference's src/main.cpp,
tests/test_small_case.cpp and tests/smoke_test.cpp with the reference's own
headers and config/result sources, unmodified, and links them against
libgsc_hip.so (outputs in oracle/_ref/, built only where the reference tree
exists; the binaries travel with the tree).
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gsc_amd import cases

REF = "/root/reference"
PKG = os.path.join(ROOT, "gpu-inner-outer_amd")
HIPLIB = os.path.join(PKG, "libgsc_hip.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")
REF_CLI = os.path.join(REF_DIR, "cuda_symbolic_control_cli_hip")
REF_SMALL = os.path.join(REF_DIR, "test_small_case_hip")
REF_SMOKE = os.path.join(REF_DIR, "smoke_test_hip")
BOUNDARY = {"gsc::cuda_backend_compiled()", "gsc::run_case_cuda_u32(gsc::CaseConfig const&)"}


def _nm(path, *flags):
    """Demangled symbol names (which may contain spaces) of `nm -C` lines
    '[address] TYPE name'."""
    out = subprocess.run(["nm", "-C", *flags, path], capture_output=True, text=True, check=True).stdout
    names = set()
    for line in out.splitlines():
        m = re.match(r"^\s*(?:[0-9a-fA-F]+\s+)?[A-Za-z]\s+(.*)$", line)
        if m:
            names.add(m.group(1))
    return names


def test_backend_exports_exactly_the_reference_symbols():
    if not os.path.exists(HIPLIB):
        pytest.fail(f"{HIPLIB} not built (make -C gpu-inner-outer_amd)")
    exported = {s for s in _nm(HIPLIB, "-D", "--defined-only") if s.startswith("gsc::")}
    assert BOUNDARY <= exported
    # The one extension entry point; nothing of the config or result code, which
    # a reference-built program brings itself.
    extra = exported - BOUNDARY
    assert extra == {"gsc::run_case_hip_u32(gsc::CaseConfig const&, gsc::CaseExtensions const&, "
                     "gsc::RunOptions const&)"}, extra


def test_struct_layouts_match_the_reference(tmp_path):
    """sizeof/offsetof of every boundary type, reference headers against the mirror."""
    if not os.path.isdir(os.path.join(REF, "include", "gsc")):
        pytest.skip("reference headers not present")
    probe = os.path.join(ROOT, "tests", "cpp", "layout_probe.cpp")
    outs = []
    for inc in (os.path.join(REF, "include"), os.path.join(PKG, "host")):
        exe = tmp_path / ("probe_" + str(len(outs)))
        subprocess.run(["g++", "-std=c++17", "-Wno-invalid-offsetof", "-I", inc, probe, "-o", str(exe)], check=True,
                       capture_output=True)
        outs.append(subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout)
    assert outs[0] == outs[1]
    assert "sizeof CaseConfig" in outs[0] and "offsetof GpuRunReport.abstraction_valid" in outs[0]


def test_reference_programs_link_against_backend():
    """Compiles (where the reference exists) and checks that each reference program
    needs nothing of the gsc namespace from outside but the backend symbols, and
    resolves them here."""
    if os.path.isdir(os.path.join(REF, "src")):
        subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "dropin"], check=True, capture_output=True)
    if not all(os.path.exists(p) for p in (REF_CLI, REF_SMALL, REF_SMOKE)):
        pytest.skip("reference programs not built (no reference tree)")
    exported = _nm(HIPLIB, "-D", "--defined-only")
    for exe in (REF_CLI, REF_SMALL, REF_SMOKE):
        needed = {s for s in _nm(exe, "--undefined-only") if s.startswith("gsc::")}
        assert "gsc::run_case_cuda_u32(gsc::CaseConfig const&)" in needed and needed <= BOUNDARY, (exe, needed)
        assert needed <= exported
        ldd = subprocess.run(["ldd", exe], capture_output=True, text=True, check=True).stdout
        assert "not found" not in ldd, ldd
        assert "libgsc_hip.so" in ldd


def _ref_program(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built (make -C oracle dropin, where the reference tree exists)")
    return path


@pytest.mark.gpu
def test_reference_small_case_test_passes(gpu):
    """tests/test_small_case.cpp: the reference's CPU solver against this backend on
    the default and wraparound cases (same_result, test_small_case.cpp:48-56)."""
    exe = _ref_program(REF_SMALL)
    print("[case] reference test_small_case against libgsc_hip.so", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    assert "small-case correctness checks passed" in p.stdout


@pytest.mark.gpu
def test_reference_smoke_test_runs(gpu):
    exe = _ref_program(REF_SMOKE)
    print("[case] reference smoke_test against libgsc_hip.so", flush=True)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    assert "gpu abstraction :" in p.stdout and "gpu solve       :" in p.stdout


@pytest.mark.gpu
def test_reference_cli_compare_and_gpu_modes(gpu, oracle, tmp_path):
    """The reference CLI (src/main.cpp) on this backend: --compare exits 0 with
    exact_match true (main.cpp:189-211), --gpu writes its HDF5 file."""
    exe = _ref_program(REF_CLI)
    case = cases.unicycle_30x30x24x10(sample_time=0.5)
    cfg = tmp_path / "c1.cfg"
    cfg.write_text(case.reference_cfg())
    print(f"[case] reference CLI --compare {case.name}", flush=True)
    p = subprocess.run([exe, str(cfg), "--compare", "--output", str(tmp_path / "cmp.h5")], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    assert "exact_match   : true" in p.stdout
    print(f"[case] reference CLI --gpu {case.name}", flush=True)
    p = subprocess.run([exe, str(cfg), "--gpu", "--output", str(tmp_path / "gpu.h5")], capture_output=True,
                       text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout[-2000:]
    ref = oracle.run(case, keep_abstraction=False)
    assert f"reachable     : {int(ref['reachable_mask'].sum())}" in p.stdout
    assert f"iterations    : {ref['iterations']}" in p.stdout
    assert (tmp_path / "gpu.h5").exists()
    h5dump = shutil.which("h5dump") or "/opt/conda/bin/h5dump"
    if os.path.exists(h5dump):
        raw = tmp_path / "reach.bin"
        subprocess.run([h5dump, "-d", "/result/gpu/reachable_mask", "-b", "LE", "-o", str(raw),
                        str(tmp_path / "gpu.h5")], capture_output=True, check=True)
        assert np.array_equal(np.fromfile(raw, np.uint8), ref["reachable_mask"])
