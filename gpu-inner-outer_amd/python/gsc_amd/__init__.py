"""gsc_amd — MI355X (gfx950) backend of the grid symbolic-control solver.

Python view of the C ABI in include/gsc_amd.h (libgsc_amd.so). All
computation happens in the HIP library; this package only describes cases,
moves host buffers and reads reports.
"""
from . import capi
from .cases import Case, builtin, load_cfg, parse_cfg
from .solver import GscError, RunResult, Solver, run_case_u32, shape_of

__all__ = ["capi", "Case", "builtin", "load_cfg", "parse_cfg", "GscError", "RunResult", "Solver", "run_case_u32",
           "shape_of"]
