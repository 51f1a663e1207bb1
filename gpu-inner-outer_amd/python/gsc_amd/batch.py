"""Many cases advanced together on one device (gsc_batch_* entry points).

Every stage of an iteration is one launch over the work of all unfinished
cases, and a fixed number of iterations is replayed from a captured HIP graph
until each case has reached its own stop rule. Results per case are identical
to a single-case solve (`Solver`) of the same case.
"""
from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Sequence

from . import capi
from .cases import Case
from .solver import GscError, RunResult, _make_buffers, _result_from, _struct_dict


def default_options() -> Dict:
    o = capi.BatchOptions()
    capi.load().gsc_batch_default_options(ctypes.byref(o))
    return _struct_dict(o)


class Batch:
    def __init__(self, cases: Sequence[Case], device: int = 0, **options):
        self.lib = capi.load()
        self.cases = list(cases)
        opts = capi.BatchOptions()
        self.lib.gsc_batch_default_options(ctypes.byref(opts))
        opts.device = device
        for k, v in options.items():
            setattr(opts, k, int(v))
        descs = [c.to_desc() for c in self.cases]  # (desc, keep-alive) pairs
        self._keep = descs
        arr = (ctypes.POINTER(capi.CaseDesc) * len(descs))(*[ctypes.pointer(d) for d, _ in descs])
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.gsc_batch_create(arr, len(descs), ctypes.byref(opts), ctypes.byref(h), err, len(err))
        if rc != capi.GSC_OK:
            raise GscError(rc, err.value.decode(errors="replace"))
        self.handle = h

    def _check(self, rc):
        if rc != capi.GSC_OK:
            raise GscError(rc, self.lib.gsc_batch_last_error(self.handle).decode(errors="replace"))

    @property
    def options(self) -> Dict:
        o = capi.BatchOptions()
        self._check(self.lib.gsc_batch_get_options(self.handle, ctypes.byref(o)))
        return _struct_dict(o)

    def set_options(self, **changes) -> None:
        o = capi.BatchOptions()
        self._check(self.lib.gsc_batch_get_options(self.handle, ctypes.byref(o)))
        for k, v in changes.items():
            setattr(o, k, int(v))
        self._check(self.lib.gsc_batch_set_options(self.handle, ctypes.byref(o)))

    def build_abstraction(self) -> float:
        ms = ctypes.c_double()
        self._check(self.lib.gsc_batch_build_abstraction(self.handle, ctypes.byref(ms)))
        return ms.value

    def solve(self) -> Dict:
        s = capi.BatchSummary()
        self._check(self.lib.gsc_batch_solve(self.handle, ctypes.byref(s)))
        return _struct_dict(s)

    def tune(self) -> List[Dict]:
        cap = 64
        entries = (capi.BatchTuneEntry * cap)()
        n = ctypes.c_uint32()
        self._check(self.lib.gsc_batch_tune(self.handle, entries, cap, ctypes.byref(n)))
        return [_struct_dict(entries[k]) for k in range(n.value)]

    def case_summary(self, index: int) -> Dict:
        s = capi.SolveSummary()
        self._check(self.lib.gsc_batch_case_summary(self.handle, index, ctypes.byref(s)))
        out = _struct_dict(s)
        out["message"] = s.message.decode(errors="replace")
        return out

    def download(self, index: int, abstraction: bool = False) -> RunResult:
        c = self.cases[index]
        S = c.state_count()
        P = S * c.input_count()
        b, arrs, stats = _make_buffers(S, P, c.max_iterations, abstraction)
        rep = capi.RunReport()
        self._check(self.lib.gsc_batch_download(self.handle, index, ctypes.byref(b), ctypes.byref(rep)))
        return _result_from(rep, arrs, stats)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gsc_batch_destroy(self.handle)
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
