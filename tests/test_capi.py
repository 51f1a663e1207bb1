"""C ABI surface and host logic (no GPU needed).

- libgsc_amd.so loads and exports every entry point declared in include/gsc_amd.h;
- the .cfg reader keeps the reference's keys, defaults, traps and error texts
  (src/config.cpp:243-347), checked against the reference build where present;
- grid shapes and their errors (grid.hpp:103-168).
"""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import ROOT
from gsc_amd import capi, cases
from gsc_amd.cases import PI_LITERAL, parse_cfg
from gsc_amd.solver import GscError, shape_of

HEADER = os.path.join(ROOT, "include", "gsc_amd.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gsc_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), f"libgsc_amd.so does not export {n}"
    # and the Python mirror binds exactly those
    assert sorted(n for n, _, _ in capi.SYMBOLS) == names


def test_struct_sizes_match_c_layout():
    # Fixed-layout PODs: a mismatch would corrupt every call.
    assert ctypes.sizeof(capi.Rect4) == 64
    assert ctypes.sizeof(capi.IterationStats) == 8 + 4 * 8 + 4 * 8
    assert ctypes.sizeof(capi.KernelTimings) == 15 * 8
    assert ctypes.sizeof(capi.MemoryBudget) == 13 * 8


C_NAMES = {
    "Rect4": "gsc_rect4", "CaseDesc": "gsc_case_desc", "CaseShape": "gsc_case_shape",
    "IterationStats": "gsc_iteration_stats", "KernelTimings": "gsc_kernel_timings",
    "MemoryStats": "gsc_memory_stats", "MemoryBudget": "gsc_memory_budget", "RunReport": "gsc_run_report",
    "ResultBuffers": "gsc_result_buffers", "SolverOptions": "gsc_solver_options",
    "SolveSummary": "gsc_solve_summary", "BatchOptions": "gsc_batch_options", "BatchSummary": "gsc_batch_summary",
    "BatchTuneEntry": "gsc_batch_tune_entry", "ServerOptions": "gsc_server_options",
    "RequestInfo": "gsc_request_info", "ServerStats": "gsc_server_stats",
}


def test_every_struct_layout_matches_the_header(tmp_path):
    """sizeof and every field offset of each ctypes mirror against the C
    compiler's view of include/gsc_amd.h."""
    structs = {n: getattr(capi, n) for n in C_NAMES}
    missing = [n for n, v in vars(capi).items()
               if isinstance(v, type) and issubclass(v, ctypes.Structure) and n not in C_NAMES]
    assert not missing, f"structs without a C name: {missing}"
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "gsc_amd.h"', "int main(void) {"]
    for py, c in C_NAMES.items():
        lines.append(f'printf("{py} %zu\\n", sizeof({c}));')
        for f, _ in structs[py]._fields_:
            lines.append(f'printf("{py}.{f} %zu\\n", offsetof({c}, {f}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = dict(line.rsplit(" ", 1) for line in out.splitlines())
    for py, cls in structs.items():
        assert int(got[py]) == ctypes.sizeof(cls), f"sizeof {py}"
        for f, _ in cls._fields_:
            assert int(got[f"{py}.{f}"]) == getattr(cls, f).offset, f"{py}.{f}"


def test_backend_compiled():
    assert capi.load().gsc_backend_compiled() == 1


PUB_CFG = """# published case
name = hyperbolic_demo
state_lb = -3.5, -3.5, -3.141592653589793, -0.1
state_ub =  3.5,  3.5,  3.141592653589793, 1.1
state_eta = 0.1, 0.1, 0.1, 0.1
wrap_dim = 2
input_lb = -1.0, -1.0
input_ub = 1.0,  1.0
input_eta = 0.2, 0.2
candidate_lb = -0.8, -0.8, -3.141592653589793, 0.0
candidate_ub =  0.8,  0.8,  3.141592653589793, 1
map_lb = -3.5, -3.5, -3.141592653589793, -0.1
map_ub =  3.5,  3.5,  3.141592653589793, 1.1
constraint_type = hyperbolic
hyperbolic_a = 4.0
hyperbolic_b = 4.0
hyperbolic_c = 16.0
disturbance_half_width = 0.05, 0.05, 0.05, 0.05
sample_time = 0.5
integration_substeps = 32
max_iterations = 100
verbose = true
"""


def test_parse_published_cfg_equals_generator():
    assert parse_cfg(PUB_CFG) == cases.hyperbolic_demo()


def test_missing_map_keeps_default_map():
    # config.cpp:244, :343-345: the map resets to the state box only when it is all zero,
    # which never happens after make_default_case(), so it stays [-2,2]^2 x ... .
    c = parse_cfg("state_lb = 0, 0, -3.141592653589793, 0\nstate_ub = 10, 10, 3.141592653589793, 2\n")
    assert c.map == ((-2.0, -2.0, -PI_LITERAL, 0.0), (2.0, 2.0, PI_LITERAL, 2.0))
    assert c.obstacles == []  # config.cpp:245 clears the default obstacle
    assert c.name == "unicycle_default"


@pytest.mark.parametrize("text,msg", [
    ("foo = 1\n", "unknown config key: foo"),
    ("state_lb = 1, 2, 3\n", "state_lb expects 4 comma-separated numbers"),
    ("input_eta = 1\n", "input_eta expects 2 comma-separated numbers"),
    ("obstacle = 1,2,3\n", "obstacle expects 8 comma-separated numbers"),
    ("constraint_type = parabolic\n", "unknown constraint_type: parabolic"),
    ("just a line\n", "expected key=value line: just a line"),
    ("sample_time = abc\n", "stod"),
])
def test_parse_errors(text, msg):
    with pytest.raises(ValueError) as ei:
        parse_cfg(text)
    assert str(ei.value) == msg


def test_comments_whitespace_and_flags():
    c = parse_cfg("  name =  spaced # trailing comment\n# full comment\n\nverbose = TRUE\nmax_iterations = 7\n"
                  "integration_substeps = 5\nwrap_dim = 3\nobstacle = 0,0,0,0, 1,1,1,1\n")
    assert c.name == "spaced" and c.verbose and c.max_iterations == 7 and c.integration_substeps == 5
    assert c.wrap_dim == 3 and c.obstacles == [((0.0,) * 4, (1.0,) * 4)]
    assert not parse_cfg("verbose = yes\n").verbose


def test_extension_keys_round_trip():
    c = cases.unicycle_91x105x42x10()
    back = parse_cfg(c.to_cfg())
    assert back == c
    b = cases.batch_instances(4, seed=9)
    for inst in b:
        assert parse_cfg(inst.to_cfg()) == inst


def test_cfg_file_load(tmp_path):
    p = tmp_path / "x.cfg"
    p.write_text(PUB_CFG)
    assert cases.load_cfg(str(p)) == cases.hyperbolic_demo()
    with pytest.raises(ValueError, match="failed to open config"):
        cases.load_cfg(str(tmp_path / "missing.cfg"))


def test_parser_matches_reference_reader(oracle, tmp_path):
    """The same file through the reference reader (oracle/_ref) and ours gives the same solve."""
    if not oracle.ref_available():
        pytest.skip("reference CPU build not present")
    text = ("name = parsed\nstate_lb = 0, 0, -3.141592653589793, 0\nstate_ub = 3,3,3.141592653589793,1\n"
            "state_eta = 0.5, 0.5, 0.7853981633974483, 0.25\ncandidate_lb = 1, 1, -4, 0\n"
            "candidate_ub=2,2,4,1\n  map_lb = 0,0,-4,0 # comment\nmap_ub = 3,3,4,1\n"
            "obstacle = 0, 2.5, -4, 0, 0.6, 3, 4, 1\nsample_time=0.5\nmax_iterations = 9\n")
    case = parse_cfg(text)
    ours = oracle.run(case)
    ref = oracle.run_ref(case)  # ref_harness rewrites the case; compare with the raw file too
    raw = tmp_path / "raw.cfg"
    raw.write_text(text)
    import subprocess
    out_dir = tmp_path / "o"
    out_dir.mkdir()
    subprocess.run([oracle.REF_HARNESS, "solve", str(raw), str(out_dir)], check=True, capture_output=True)
    for k, t in (("reachable_mask", np.uint8), ("controller", np.uint32), ("valid_mask", np.uint8)):
        assert np.array_equal(np.fromfile(out_dir / f"{k}.bin", dtype=t), ours[k]), k
        assert np.array_equal(ref[k], ours[k]), k


def test_shape_errors():
    c = cases.builtin("default")
    c.state_eta = (0.5, 0.0, 0.5, 0.5)
    with pytest.raises(GscError, match="state eta must be positive"):
        shape_of(c)
    c = cases.builtin("default")
    c.state_ub = (2.0, -3.0, PI_LITERAL, 2.0)
    with pytest.raises(GscError, match="state bounds must satisfy ub > lb"):
        shape_of(c)
    c = cases.builtin("default")
    c.input_eta = (-1.0, 1.0)
    with pytest.raises(GscError, match="input eta must be positive"):
        shape_of(c)
    c = cases.builtin("default")
    c.state_eta = (1e-4, 1e-4, 1e-3, 1e-3)
    with pytest.raises(GscError, match="state grid exceeds the selected state index type"):
        shape_of(c)
    s = shape_of(cases.hyperbolic_demo())
    assert s["state_shape"] == (70, 70, 63, 13) and s["pair_count"] == 401310000


def test_generators_are_deterministic_and_in_range():
    a = cases.batch_instances(32, seed=3)
    b = cases.batch_instances(32, seed=3)
    assert a == b
    for inst in a:
        n = inst.state_count()
        assert 3e4 <= n <= 1.3e6, n
        assert inst.input_count() in (16, 36, 64, 100)
        assert 0 <= len(inst.extra_targets) <= 2
        assert len(inst.obstacles) <= 5
    r = cases.serving_requests(16, seed=5)
    assert all(1e4 <= x.state_count() <= 5e6 for x in r)
