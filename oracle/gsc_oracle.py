"""Python bridge to the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker / baseline — never as the product path.

  liboracle.so           this project's CPU restatement of the reference
                         (oracle/gsc_oracle.cpp, file:line citations inside)
  _ref/ref_harness       the reference's own CPU solver compiled from its
                         sources (oracle/Makefile `ref` target), present only
                         where it was built

Case descriptors use the product's C struct layout (include/gsc_amd.h), via
gsc_amd.cases.Case.to_desc — a shared data format, no product computation.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess
import tempfile
from ctypes import POINTER, c_char_p, c_int, c_int32, c_size_t, c_uint32, c_uint64, c_void_p
from typing import Dict, Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_HARNESS = os.path.join(HERE, "_ref", "ref_harness")

_lib = None


def _capi():
    import sys
    pkg = os.path.join(os.path.dirname(HERE), "gpu-inner-outer_amd", "python")
    if pkg not in sys.path:
        sys.path.insert(0, pkg)
    from gsc_amd import capi  # noqa: E402
    return capi


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} missing; run `make -C {HERE}`")
    capi = _capi()
    lib = ctypes.CDLL(LIB_PATH)
    D = POINTER(capi.CaseDesc)
    u8p = POINTER(ctypes.c_uint8)
    u32p = POINTER(c_uint32)
    u64p = POINTER(c_uint64)
    lib.oracle_shape.argtypes = [D, POINTER(capi.CaseShape), c_char_p, c_size_t]
    lib.oracle_masks.argtypes = [D, u8p, u8p, c_char_p, c_size_t]
    lib.oracle_successor_box.argtypes = [D, c_uint32, c_uint32, POINTER(ctypes.c_double), POINTER(ctypes.c_double)]
    lib.oracle_abstraction.argtypes = [D, c_uint64, c_uint64, u32p, u32p, u8p, c_int]
    lib.oracle_prefix_sum.argtypes = [D, u8p, u64p]
    lib.oracle_query_box_count.argtypes = [D, u64p, u32p, u32p, c_uint64, u64p]
    lib.oracle_solve.argtypes = [D, u8p, u8p, u32p, u32p, u8p, u8p, u32p, u32p, u64p, u64p, c_uint32,
                                 u32p, POINTER(c_int32), c_char_p, c_size_t, c_int]
    for f in ("oracle_shape", "oracle_masks", "oracle_successor_box", "oracle_abstraction", "oracle_prefix_sum",
              "oracle_query_box_count", "oracle_solve"):
        getattr(lib, f).restype = c_int
    _lib = lib
    return lib


def _p(a, t):
    return a.ctypes.data_as(POINTER(t))


def shape(case) -> Dict:
    lib = load()
    capi = _capi()
    d, keep = case.to_desc()
    sh = capi.CaseShape()
    err = ctypes.create_string_buffer(256)
    rc = lib.oracle_shape(ctypes.byref(d), ctypes.byref(sh), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return {"state_shape": tuple(sh.state_shape), "state_count": sh.state_count,
            "input_shape": tuple(sh.input_shape), "input_count": sh.input_count, "pair_count": sh.pair_count,
            "prefix_count": sh.prefix_count, "prefix_shape": tuple(sh.prefix_shape)}


def masks(case):
    lib = load()
    S = shape(case)["state_count"]
    d, keep = case.to_desc()
    valid = np.zeros(S, np.uint8)
    cand = np.zeros(S, np.uint8)
    err = ctypes.create_string_buffer(256)
    rc = lib.oracle_masks(ctypes.byref(d), _p(valid, ctypes.c_uint8), _p(cand, ctypes.c_uint8), err, 256)
    if rc:
        raise ValueError(err.value.decode())
    return valid, cand


def successor_box(case, state: int, inp: int):
    lib = load()
    d, keep = case.to_desc()
    mn = (ctypes.c_double * 4)()
    mx = (ctypes.c_double * 4)()
    lib.oracle_successor_box(ctypes.byref(d), state, inp, mn, mx)
    return tuple(mn), tuple(mx)


def abstraction(case, state_begin: int = 0, state_end: Optional[int] = None, threads: int = 8):
    lib = load()
    sh = shape(case)
    S, I = sh["state_count"], sh["input_count"]
    if state_end is None or state_end > S:
        state_end = S
    n = (state_end - state_begin) * I
    mn = np.zeros(n, np.uint32)
    mx = np.zeros(n, np.uint32)
    va = np.zeros(n, np.uint8)
    d, keep = case.to_desc()
    rc = lib.oracle_abstraction(ctypes.byref(d), state_begin, state_end, _p(mn, c_uint32), _p(mx, c_uint32),
                                _p(va, ctypes.c_uint8), threads)
    if rc:
        raise ValueError("oracle_abstraction failed")
    return mn, mx, va


def prefix_sum(case, mask: np.ndarray) -> np.ndarray:
    lib = load()
    sh = shape(case)
    t = np.zeros(sh["prefix_count"], np.uint64)
    d, keep = case.to_desc()
    m = np.ascontiguousarray(mask, np.uint8)
    lib.oracle_prefix_sum(ctypes.byref(d), _p(m, ctypes.c_uint8), _p(t, c_uint64))
    return t


def query_box_count(case, table: np.ndarray, min_flat: np.ndarray, max_flat: np.ndarray) -> np.ndarray:
    lib = load()
    d, keep = case.to_desc()
    a = np.ascontiguousarray(min_flat, np.uint32)
    b = np.ascontiguousarray(max_flat, np.uint32)
    out = np.zeros(len(a), np.uint64)
    lib.oracle_query_box_count(ctypes.byref(d), _p(np.ascontiguousarray(table, np.uint64), c_uint64),
                               _p(a, c_uint32), _p(b, c_uint32), len(a), _p(out, c_uint64))
    return out


def solve(case, valid, cand, mn, mx, va, threads: int = 8) -> Dict:
    lib = load()
    S = len(valid)
    d, keep = case.to_desc()
    reach = np.zeros(S, np.uint8)
    ctrl = np.zeros(S, np.uint32)
    rstep = np.zeros(S, np.uint32)
    cap = max(1, case.max_iterations)
    nr = np.zeros(cap, np.uint64)
    nc = np.zeros(cap, np.uint64)
    iters = c_uint32()
    conv = c_int32()
    msg = ctypes.create_string_buffer(256)
    rc = lib.oracle_solve(ctypes.byref(d), _p(valid, ctypes.c_uint8), _p(cand, ctypes.c_uint8), _p(mn, c_uint32),
                          _p(mx, c_uint32), _p(va, ctypes.c_uint8), _p(reach, ctypes.c_uint8), _p(ctrl, c_uint32),
                          _p(rstep, c_uint32), _p(nr, c_uint64), _p(nc, c_uint64), cap, ctypes.byref(iters),
                          ctypes.byref(conv), msg, 256, threads)
    if rc:
        raise ValueError(msg.value.decode())
    it = iters.value
    return {"reachable_mask": reach, "controller": ctrl, "reach_step": rstep, "iterations": it,
            "converged": bool(conv.value), "message": msg.value.decode(),
            "newly_reachable": nr[:min(it, cap)].copy(), "newly_certified": nc[:min(it, cap)].copy(),
            "reachable_states": int(reach.sum()), "certified_candidate_states": int(nc[:min(it, cap)].sum())}


def run(case, threads: int = 8, keep_abstraction: bool = True) -> Dict:
    """Masks + abstraction + fixpoint, the reference `--cpu` pipeline (main.cpp:157-159)."""
    valid, cand = masks(case)
    mn, mx, va = abstraction(case, threads=threads)
    out = solve(case, valid, cand, mn, mx, va, threads=threads)
    out["valid_mask"] = valid
    out["candidate_mask"] = cand
    if keep_abstraction:
        out["min_flat"], out["max_flat"], out["pair_valid"] = mn, mx, va
    return out


# ---------------------------------------------------------------------------
# The reference's own CPU build (oracle/_ref), where it was built.
# ---------------------------------------------------------------------------
def ref_available() -> bool:
    return os.path.exists(REF_HARNESS)


def _extension_args(case):
    args = []
    for lb, ub in case.extra_targets:
        args += ["--target"] + ["%.17g" % v for v in tuple(lb) + tuple(ub)]
    if case.target_band is not None:
        args += ["--band", "%.17g" % case.target_band[0], "%.17g" % case.target_band[1]]
    return args


_ITER_RE = re.compile(r"iter (\d+): new_reachable=(\d+) candidate_controller_coverage=[\d.]+% \((\d+)/(\d+)\)")


def run_ref(case, timeout: float = 3600.0, builtin: Optional[str] = None,
            max_iterations: Optional[int] = None) -> Dict:
    """Run the reference CPU solver (load_case_config + prepare_case_cpu +
    solve_backward_cpu) on `case` written as a `.cfg` file — or, with
    `builtin`, on the reference's own canned case of that name
    (make_default_case etc., config.cpp:129-241); `case` then only sizes the
    outputs and must describe the same grid."""
    if not ref_available():
        raise RuntimeError("oracle/_ref/ref_harness not built")
    sh = shape(case)
    S, P = sh["state_count"], sh["pair_count"]
    env = dict(os.environ)
    if max_iterations is not None:
        env["REF_MAX_ITERATIONS"] = str(max_iterations)
    with tempfile.TemporaryDirectory() as tmp:
        cfg = os.path.join(tmp, "case.cfg")
        with open(cfg, "w") as f:
            f.write(case.reference_cfg().replace("verbose = false", "verbose = true"))
        src = f"builtin:{builtin}" if builtin else cfg
        ext = [] if builtin else _extension_args(case)
        p = subprocess.run([REF_HARNESS, "solve", src, tmp] + ext, capture_output=True, text=True,
                           timeout=timeout, env=env)
        if p.returncode != 0:
            raise RuntimeError(p.stderr.strip() or p.stdout[-2000:])
        rd = lambda n, t: np.fromfile(os.path.join(tmp, n), dtype=t)
        out = {
            "valid_mask": rd("valid_mask.bin", np.uint8), "candidate_mask": rd("candidate_mask.bin", np.uint8),
            "reachable_mask": rd("reachable_mask.bin", np.uint8), "controller": rd("controller.bin", np.uint32),
            "reach_step": rd("reach_step.bin", np.uint32), "min_flat": rd("min_flat.bin", np.uint32),
            "max_flat": rd("max_flat.bin", np.uint32), "pair_valid": rd("pair_valid.bin", np.uint8),
        }
        summ = {}
        with open(os.path.join(tmp, "summary.txt")) as f:
            for line in f:
                k, _, v = line.strip().partition(" ")
                summ[k] = v
    assert len(out["valid_mask"]) == S and len(out["min_flat"]) == P
    nr, cert = [], []
    for m in _ITER_RE.finditer(p.stdout):
        nr.append(int(m.group(2)))
        cert.append(int(m.group(3)))
    nc = [cert[0]] + [cert[k] - cert[k - 1] for k in range(1, len(cert))] if cert else []
    out.update({
        "iterations": int(summ["iterations"]), "converged": summ["converged"] == "1",
        "message": summ.get("message", ""), "reachable_states": int(summ["reachable_states"]),
        "certified_candidate_states": int(summ["certified_candidate_states"]),
        "abstraction_ms": float(summ["abstraction_ms"]), "solve_ms": float(summ["solve_ms"]),
        "newly_reachable": np.array(nr, np.uint64), "newly_certified": np.array(nc, np.uint64),
    })
    return out


def ref_sample(case, state_begin: int, state_end: int, sweeps: int, timeout: float = 600.0) -> Dict:
    """Time the reference's per-pair work on a slice of states (see ref_harness.cpp)."""
    if not ref_available():
        raise RuntimeError("oracle/_ref/ref_harness not built")
    with tempfile.TemporaryDirectory() as tmp:
        cfg = os.path.join(tmp, "case.cfg")
        with open(cfg, "w") as f:
            f.write(case.reference_cfg().replace("verbose = true", "verbose = false"))
        p = subprocess.run([REF_HARNESS, "sample", cfg, str(state_begin), str(state_end), str(sweeps)]
                           + _extension_args(case), capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0:
            raise RuntimeError(p.stderr.strip())
    out = {}
    for line in p.stdout.splitlines():
        k, _, v = line.partition(" ")
        if k in ("sample_states", "sample_pairs", "abstraction_ms", "satisfaction_sweeps", "satisfaction_ms",
                 "satisfied"):
            out[k] = float(v)
    return out
