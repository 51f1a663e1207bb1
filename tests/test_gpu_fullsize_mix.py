"""Batching and serving at their full workload sizes against the reference.

Every case of the 32-case batching workload (batch_instances(32, seed=3):
5e4-1e6 cells, 16-100 inputs, 535M pairs) and of the 48-request serving
workload (serving_requests(48, seed=5): 2e4-4e6 cells, eight requests above
3M cells, 2.8G pairs) was solved on its own by the reference's CPU build
(tests/golden/make_goldens.py --batch / --serving). Here the batched path
(default options, then autotuned) and the server (Poisson arrivals, many
requests in flight) must return, for each case, arrays with the reference's
digests and the reference's per-iteration counts: the same-result rule of
test_small_case.cpp:48-56 applied per instance.
"""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gsc_amd import cases
from gsc_amd.batch import Batch
from gsc_amd.serve import Server, poisson_load

pytestmark = pytest.mark.gpu

ARRAYS = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step")


def _golden(name):
    path = os.path.join(GOLDEN, f"{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"golden {name} not generated")
    with open(path) as f:
        return json.load(f)


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _assert_reference(got, fx, label):
    assert got.executed, f"{label}: {got.message}"
    for k in ARRAYS:
        assert _sha(getattr(got, k)) == fx["sha256"][k], f"{label}: {k} digest differs from the reference"
    assert got.iterations == fx["iterations"], f"{label}: iterations {got.iterations} vs {fx['iterations']}"
    assert got.result_message == fx["message"], label
    assert [s["newly_reachable"] for s in got.iteration_stats] == fx["newly_reachable"], label
    assert [s["newly_certified"] for s in got.iteration_stats] == fx["newly_certified"], label
    assert got.reachable_states == fx["reachable_states"], label
    assert got.certified_candidate_states == fx["certified_candidate_states"], label


def _cases_for(fixtures, generated):
    # The generator must reproduce the exact configurations the goldens were made from.
    for c in generated:
        assert c.to_cfg() == fixtures[c.name]["cfg"], f"{c.name}: generated case differs from the golden's"
    return generated


def test_batch_full_size_matches_reference(gpu):
    fx = _golden("batch_instances_32_s3")
    cs = _cases_for(fx, cases.batch_instances(32, seed=3))
    print(f"[case] full-size batch: {len(cs)} cases, {sum(c.state_count() * c.input_count() for c in cs)} pairs",
          flush=True)
    with Batch(cs) as b:
        b.build_abstraction()
        summ = b.solve()
        assert summ["stopped"] == len(cs)
        for k, c in enumerate(cs):
            _assert_reference(b.download(k), fx[c.name], f"batch default {c.name}")
        print("[case] full-size batch, tuned", flush=True)
        b.tune()
        b.build_abstraction()
        b.solve()
        for k, c in enumerate(cs):
            _assert_reference(b.download(k), fx[c.name], f"batch tuned {c.name}")


def test_serving_full_size_matches_reference(gpu):
    fx = _golden("serving_requests_48_s5")
    reqs = _cases_for(fx, cases.serving_requests(48, seed=5))
    assert sum(c.state_count() >= 3_000_000 for c in reqs) >= 1
    print(f"[case] full-size serving: {len(reqs)} requests at 400/s", flush=True)
    with Server(max_active=64, iterations_per_step=4, hbm_budget_bytes=96 << 30) as srv:
        stats, results = poisson_load(srv, reqs, rate_per_s=400.0, seed=400, keep_results=True)
        assert not stats["failed"], stats["failed"]
        assert stats["server"]["max_active_seen"] > 1  # requests really overlapped on the device
        for c, r in zip(reqs, results):
            _assert_reference(r, fx[c.name], f"served {c.name}")
