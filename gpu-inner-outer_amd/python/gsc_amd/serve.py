"""Serving: a request loop on one device with continuous batching (gsc_server_*).

`Server.submit` returns at once; a worker thread in the library admits cases
against an HBM budget, prepares them while earlier ones iterate, advances all
admitted cases together and hands results back as each one stops. Results are
identical to a single-case solve of the same case.

`poisson_load` drives a server with a Poisson arrival stream and reports
per-request latency; `sequential_load` serves the same stream one request at a
time through the plain per-call path (gsc_run_case_u32), as a baseline.
"""
from __future__ import annotations

import ctypes
import threading
import time
from typing import Dict, List, Optional, Sequence

import numpy as np

from . import capi
from .cases import Case
from .solver import GscError, RunResult, _make_buffers, _result_from, _struct_dict, run_case_u32


class Server:
    def __init__(self, device: int = 0, max_active: int = 64, iterations_per_step: int = 4,
                 hbm_budget_bytes: int = 64 << 30, tile_wg: int = 512, tile_ib: int = 8):
        self.lib = capi.load()
        o = capi.ServerOptions()
        self.lib.gsc_server_default_options(ctypes.byref(o))
        o.device, o.max_active, o.iterations_per_step = device, max_active, iterations_per_step
        o.hbm_budget_bytes, o.tile_wg, o.tile_ib = hbm_budget_bytes, tile_wg, tile_ib
        h = ctypes.c_void_p()
        err = ctypes.create_string_buffer(512)
        rc = self.lib.gsc_server_create(ctypes.byref(o), ctypes.byref(h), err, len(err))
        if rc != capi.GSC_OK:
            raise GscError(rc, err.value.decode(errors="replace"))
        self.handle = h
        self._cases: Dict[int, Case] = {}

    def submit(self, case: Case) -> int:
        d, keep = case.to_desc()
        rid = ctypes.c_uint64()
        rc = self.lib.gsc_server_submit(self.handle, ctypes.byref(d), ctypes.byref(rid))
        if rc != capi.GSC_OK:
            raise GscError(rc, "server is not accepting requests")
        self._cases[rid.value] = case
        return rid.value

    def wait(self, rid: int, timeout_ms: float = -1.0) -> Optional[Dict]:
        """Request info once finished (done or failed); None on timeout."""
        info = capi.RequestInfo()
        rc = self.lib.gsc_server_wait(self.handle, rid, timeout_ms, ctypes.byref(info))
        if rc == capi.GSC_ERR_BUSY:
            return None
        if rc == capi.GSC_ERR_INVALID_ARGUMENT:
            raise GscError(rc, f"unknown request {rid}")
        out = _struct_dict(info)
        out["summary"] = _struct_dict(info.summary)
        out["summary"]["message"] = info.summary.message.decode(errors="replace")
        out["error"] = info.error.decode(errors="replace")
        return out

    def fetch(self, rid: int) -> RunResult:
        """Results of a finished request (raises for a failed one); forgets it."""
        c = self._cases.pop(rid)
        S = c.state_count()
        b, arrs, stats = _make_buffers(S, 0, c.max_iterations, False)
        rep = capi.RunReport()
        rc = self.lib.gsc_server_fetch(self.handle, rid, ctypes.byref(b), ctypes.byref(rep))
        if rc != capi.GSC_OK:
            raise GscError(rc, rep.message.decode(errors="replace"))
        return _result_from(rep, arrs, stats)

    def stats(self) -> Dict:
        s = capi.ServerStats()
        self.lib.gsc_server_get_stats(self.handle, ctypes.byref(s))
        return _struct_dict(s)

    def close(self):
        if getattr(self, "handle", None):
            self.lib.gsc_server_destroy(self.handle)
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


def arrival_times(n: int, rate_per_s: float, seed: int) -> np.ndarray:
    """Poisson process: exponential inter-arrival gaps with mean 1/rate."""
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_per_s, size=n))


def _summarise(lat_ms: Sequence[float], span_s: float, n: int) -> Dict:
    lat = np.asarray(lat_ms, dtype=np.float64)
    return {
        "requests": n,
        "solves_per_s": n / span_s,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "mean_ms": float(lat.mean()),
        "max_ms": float(lat.max()),
        "span_s": span_s,
    }


def poisson_load(server: Server, cases: Sequence[Case], rate_per_s: float, seed: int = 0,
                 keep_results: bool = False):
    """Submit `cases` at Poisson arrival times; latency = submit -> results in the
    server's host memory (measured by the server). The span runs from the start
    of the stream to the last completion (submit time + latency), so copying the
    results out to Python afterwards is not part of the served throughput."""
    arrivals = arrival_times(len(cases), rate_per_s, seed)
    ids: List[int] = [0] * len(cases)
    submitted_at: List[float] = [0.0] * len(cases)
    t0 = time.perf_counter()

    def submitter():
        for k, c in enumerate(cases):
            delay = arrivals[k] - (time.perf_counter() - t0)
            if delay > 0:
                time.sleep(delay)
            submitted_at[k] = time.perf_counter() - t0
            ids[k] = server.submit(c)

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    lat, results, failed, done_at = [], [], [], []
    for k, rid in enumerate(ids):
        info = server.wait(rid)
        if info["state"] != capi.REQUEST_DONE:
            failed.append((k, info["error"]))
            server._cases.pop(rid, None)
            server.lib.gsc_server_fetch(server.handle, rid, None, None)
            continue
        lat.append(info["latency_ms"])
        done_at.append(submitted_at[k] + info["latency_ms"] / 1000.0)
        if keep_results:
            results.append(server.fetch(rid))
        else:  # release the request without copying its arrays
            server._cases.pop(rid, None)
            server.lib.gsc_server_fetch(server.handle, rid, None, None)
            results.append(None)
    span = max(done_at) if done_at else time.perf_counter() - t0
    out = _summarise(lat, span, len(lat))
    out["rate_per_s"] = rate_per_s
    out["failed"] = failed
    out["server"] = server.stats()
    return out, results


def sequential_load(cases: Sequence[Case], rate_per_s: float, seed: int = 0, keep_results: bool = False):
    """The same arrival stream served one request at a time by the plain per-call
    path (gsc_run_case_u32: allocate, abstraction, fixpoint, copy back, free);
    a request waits for the ones before it."""
    arrivals = arrival_times(len(cases), rate_per_s, seed)
    t0 = time.perf_counter()
    lat, results = [], []
    for k, c in enumerate(cases):
        now = time.perf_counter() - t0
        if arrivals[k] > now:
            time.sleep(arrivals[k] - now)
        r = run_case_u32(c)
        done = time.perf_counter() - t0
        lat.append((done - arrivals[k]) * 1000.0)
        results.append(r if keep_results else None)
    span = time.perf_counter() - t0
    out = _summarise(lat, span, len(cases))
    out["rate_per_s"] = rate_per_s
    return out, results
