"""Python-side mirror of the backend entry points, over the C ABI.

`run_case_u32(case)` is the counterpart of the reference's
`gsc::run_case_cuda_u32(const CaseConfig&)` (include/gsc/cuda_api.hpp:88):
it returns the masks, controller and reach-step tables, per-iteration
statistics and timings, and — like the reference — raises for preparation
errors (kernels.cu:505 lets them escape) while device errors are reported in
`report.message` with `executed == False`.

`Solver` keeps one case resident on the device across calls (abstraction and
iteration state stay in HBM), which the bench and the batch/serve layers use.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import capi
from .cases import Case


class GscError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(message)
        self.status = status


def _struct_dict(s) -> Dict:
    return {name: getattr(s, name) for name, _ in s._fields_ if name != "reserved"}


def shape_of(case: Case) -> Dict:
    lib = capi.load()
    d, keep = case.to_desc()
    sh = capi.CaseShape()
    err = ctypes.create_string_buffer(512)
    rc = lib.gsc_case_shape_of(ctypes.byref(d), ctypes.byref(sh), err, len(err))
    if rc != capi.GSC_OK:
        raise GscError(rc, err.value.decode())
    return {
        "state_shape": tuple(sh.state_shape),
        "state_count": sh.state_count,
        "input_shape": tuple(sh.input_shape),
        "input_count": sh.input_count,
        "pair_count": sh.pair_count,
        "prefix_shape": tuple(sh.prefix_shape),
        "prefix_count": sh.prefix_count,
    }


@dataclass
class RunResult:
    executed: bool
    message: str
    result_message: str
    converged: bool
    iterations: int
    reachable_states: int
    certified_candidate_states: int
    candidate_states: int
    pair_count: int
    abstraction_ms: float
    solve_ms: float
    total_ms: float
    valid_mask: Optional[np.ndarray] = None
    candidate_mask: Optional[np.ndarray] = None
    reachable_mask: Optional[np.ndarray] = None
    controller: Optional[np.ndarray] = None
    reach_step: Optional[np.ndarray] = None
    abstraction_min_flat: Optional[np.ndarray] = None
    abstraction_max_flat: Optional[np.ndarray] = None
    abstraction_valid: Optional[np.ndarray] = None
    iteration_stats: List[Dict] = field(default_factory=list)
    kernel_timings: Dict = field(default_factory=dict)
    memory_stats: Dict = field(default_factory=dict)
    budget: Dict = field(default_factory=dict)
    shape: Dict = field(default_factory=dict)


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _make_buffers(S: int, P: int, max_iter: int, abstraction: bool):
    arrs = {
        "valid_mask": np.zeros(S, np.uint8),
        "candidate_mask": np.zeros(S, np.uint8),
        "reachable_mask": np.zeros(S, np.uint8),
        "controller": np.zeros(S, np.uint32),
        "reach_step": np.zeros(S, np.uint32),
    }
    if abstraction:
        arrs["abstraction_min_flat"] = np.zeros(P, np.uint32)
        arrs["abstraction_max_flat"] = np.zeros(P, np.uint32)
        arrs["abstraction_valid"] = np.zeros(P, np.uint8)
    cap = max(1, min(max_iter, 1 << 20))
    stats = (capi.IterationStats * cap)()
    b = capi.ResultBuffers()
    for k, a in arrs.items():
        setattr(b, k, _ptr(a))
    b.iteration_stats = ctypes.cast(stats, ctypes.POINTER(capi.IterationStats))
    b.iteration_stats_capacity = cap
    return b, arrs, stats


def _result_from(report: capi.RunReport, arrs, stats) -> RunResult:
    n = report.iteration_stats_written
    r = RunResult(
        executed=bool(report.executed),
        message=report.message.decode(errors="replace"),
        result_message=report.result_message.decode(errors="replace"),
        converged=bool(report.converged),
        iterations=report.iterations,
        reachable_states=report.reachable_states,
        certified_candidate_states=report.certified_candidate_states,
        candidate_states=report.candidate_states,
        pair_count=report.pair_count,
        abstraction_ms=report.abstraction_ms,
        solve_ms=report.solve_ms,
        total_ms=report.total_ms,
        iteration_stats=[_struct_dict(stats[k]) for k in range(n)],
        kernel_timings=_struct_dict(report.kernel_timings),
        memory_stats=_struct_dict(report.memory_stats),
        budget=_struct_dict(report.budget),
        shape={"state_shape": tuple(report.shape.state_shape), "input_shape": tuple(report.shape.input_shape)},
    )
    for k, a in arrs.items():
        setattr(r, k, a)
    return r


def run_case_u32(case: Case, abstraction: bool = False) -> RunResult:
    """gsc_run_case_u32 — the drop-in `run_case_cuda_u32` (kernels.cu:492-875)."""
    lib = capi.load()
    sh = shape_of(case)  # raises GscError for the reference's grid errors
    d, keep = case.to_desc()
    b, arrs, stats = _make_buffers(sh["state_count"], sh["pair_count"], case.max_iterations, abstraction)
    rep = capi.RunReport()
    rc = lib.gsc_run_case_u32(ctypes.byref(d), ctypes.byref(b), ctypes.byref(rep))
    if rc in (capi.GSC_ERR_CONFIG, capi.GSC_ERR_INVALID_ARGUMENT):
        raise GscError(rc, rep.message.decode(errors="replace"))
    return _result_from(rep, arrs, stats)


def run_case_u32_static(case: Case, abstraction: bool = False) -> RunResult:
    """gsc_run_case_u32_static — run_case_u32 with every device buffer taken from
    the library's fixed device array (no allocation). Raises GscError for
    configuration errors and for cases larger than the array."""
    lib = capi.load()
    sh = shape_of(case)
    d, keep = case.to_desc()
    b, arrs, stats = _make_buffers(sh["state_count"], sh["pair_count"], case.max_iterations, abstraction)
    rep = capi.RunReport()
    rc = lib.gsc_run_case_u32_static(ctypes.byref(d), ctypes.byref(b), ctypes.byref(rep))
    if rc in (capi.GSC_ERR_CONFIG, capi.GSC_ERR_INVALID_ARGUMENT, capi.GSC_ERR_OUT_OF_MEMORY):
        raise GscError(rc, rep.message.decode(errors="replace"))
    return _result_from(rep, arrs, stats)


class Solver:
    """One case resident on one device (gsc_solver_* entry points)."""

    def __init__(self, case: Case, device: int = 0, keep_export: bool = False, time_kernels: bool = True,
                 sweep_all_pairs: bool = False, verbose: bool = False, recompute_abstraction: bool = False,
                 tile_wg: int = 0, tile_ib: int = 0):
        self.lib = capi.load()
        self.case = case
        opts = capi.SolverOptions()
        self.lib.gsc_solver_default_options(ctypes.byref(opts))
        opts.device = device
        opts.keep_export = int(keep_export)
        opts.time_kernels = int(time_kernels)
        opts.sweep_all_pairs = int(sweep_all_pairs)
        opts.verbose = int(verbose)
        opts.recompute_abstraction = int(recompute_abstraction)
        opts.tile_wg, opts.tile_ib = int(tile_wg), int(tile_ib)
        d, keep = case.to_desc()
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.gsc_solver_create(ctypes.byref(d), ctypes.byref(opts), ctypes.byref(h), err, len(err))
        if rc != capi.GSC_OK:
            raise GscError(rc, err.value.decode(errors="replace"))
        self.handle = h
        sh = capi.CaseShape()
        self.lib.gsc_solver_shape(self.handle, ctypes.byref(sh))
        self.state_count = sh.state_count
        self.pair_count = sh.pair_count
        self.input_count = sh.input_count
        self.state_shape = tuple(sh.state_shape)

    def _check(self, rc):
        if rc != capi.GSC_OK:
            raise GscError(rc, self.lib.gsc_solver_last_error(self.handle).decode(errors="replace"))

    def build_abstraction(self) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.gsc_solver_build_abstraction(self.handle, ctypes.byref(ms)))
        return ms.value

    def solve(self) -> Dict:
        s = capi.SolveSummary()
        self._check(self.lib.gsc_solver_solve(self.handle, ctypes.byref(s)))
        out = _struct_dict(s)
        out["message"] = s.message.decode(errors="replace")
        return out

    def download(self, abstraction: bool = False) -> RunResult:
        b, arrs, stats = _make_buffers(self.state_count, self.pair_count, self.case.max_iterations, abstraction)
        rep = capi.RunReport()
        self._check(self.lib.gsc_solver_download(self.handle, ctypes.byref(b), ctypes.byref(rep)))
        return _result_from(rep, arrs, stats)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gsc_solver_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
