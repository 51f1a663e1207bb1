"""Full-size GPU parity: 401,310,000-pair cases against golden digests written
by the reference's own CPU build (tests/golden/make_goldens.py --large)."""
import hashlib
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN
from gsc_amd import cases
from gsc_amd.solver import run_case_u32

pytestmark = pytest.mark.gpu


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def _golden(name):
    path = os.path.join(GOLDEN, f"{name}.json")
    if not os.path.exists(path):
        pytest.skip(f"golden {name} not generated")
    with open(path) as f:
        return json.load(f)


def _check(case, fx, abstraction):
    print(f"[case] {case.name} pairs={case.state_count() * case.input_count()}", flush=True)
    got = run_case_u32(case, abstraction=abstraction)
    assert got.executed, got.message
    names = {"valid_mask": "valid_mask", "candidate_mask": "candidate_mask", "reachable_mask": "reachable_mask",
             "controller": "controller", "reach_step": "reach_step"}
    if abstraction:
        names.update({"abstraction_min_flat": "min_flat", "abstraction_max_flat": "max_flat",
                      "abstraction_valid": "pair_valid"})
    for attr, key in names.items():
        assert sha(getattr(got, attr)) == fx["sha256"][key], f"{attr} digest differs"
    assert got.iterations == fx["iterations"]
    assert got.result_message == fx["message"]
    assert [s["newly_reachable"] for s in got.iteration_stats] == fx["newly_reachable"]
    assert [s["newly_certified"] for s in got.iteration_stats] == fx["newly_certified"]
    assert got.reachable_states == fx["reachable_states"]
    assert got.certified_candidate_states == fx["certified_candidate_states"]
    return got


def test_published_case_bit_exact(gpu):
    """examples/hyperbolic_case.cfg: 70x70x63x13 x 100 (README.md:44-98)."""
    got = _check(cases.hyperbolic_demo(), _golden("hyperbolic_demo"), abstraction=True)
    assert got.reachable_states == 1142689 and got.iterations == 12 and got.converged


def test_unicycle_91x105x42x10_bit_exact(gpu):
    _check(cases.unicycle_91x105x42x10(), _golden("unicycle_91x105x42x10"), abstraction=True)


def test_unicycle_128x128x64x16_certified(gpu, oracle):
    """1,677,721,600 pairs: too large for a full oracle run in a test, so the
    result is certified on sampled states against the oracle's own successor
    boxes and queries (tests/certify.py), and the three device schedules —
    stored records, recomputed boxes, every pair every iteration — must agree
    bit for bit."""
    from certify import certify
    from casegen import SOLVE_KEYS
    from gsc_amd.solver import Solver

    case = cases.unicycle_128x128x64x16()
    print(f"[case] {case.name} stored", flush=True)
    with Solver(case, time_kernels=False) as s:
        s.build_abstraction()
        s.solve()
        stored = s.download()
    out = certify(oracle, case, stored, n_chunks=40, chunk=64, seed=1)
    assert out["checked"] > 1000
    for label, kw in (("recompute", {"recompute_abstraction": True}), ("sweep", {"sweep_all_pairs": True})):
        print(f"[case] {case.name} {label}", flush=True)
        with Solver(case, time_kernels=False, **kw) as s:
            s.build_abstraction()
            s.solve()
            other = s.download()
        for k in SOLVE_KEYS:
            assert np.array_equal(getattr(other, k), getattr(stored, k)), f"{label}: {k}"
        assert other.iterations == stored.iterations and other.result_message == stored.result_message
        assert [a["newly_reachable"] for a in other.iteration_stats] == \
            [a["newly_reachable"] for a in stored.iteration_stats]
