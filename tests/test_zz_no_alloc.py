"""gsc_run_case_u32_static (no allocation) against gsc_run_case_u32, through the C ABI.

Two calls in a row share the library's device array; the second case is
smaller, so anything the second run read without first writing it would still
hold the first run's data.
"""
import numpy as np
import pytest

from gsc_amd import capi, cases
from gsc_amd.solver import GscError, run_case_u32, run_case_u32_static

KEYS = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step")
ABS = ("abstraction_min_flat", "abstraction_max_flat", "abstraction_valid")


def _same(a, b, label, abstraction):
    assert a.executed and b.executed, label
    for k in KEYS + (ABS if abstraction else ()):
        x, y = getattr(a, k), getattr(b, k)
        assert x.shape == y.shape, f"{label}: {k} shape"
        assert np.array_equal(x, y), f"{label}: {k} differs at {np.count_nonzero(x != y)} entries"
    assert a.iterations == b.iterations and a.converged == b.converged, label
    assert a.result_message == b.result_message, label
    assert a.reachable_states == b.reachable_states, label
    assert a.certified_candidate_states == b.certified_candidate_states, label
    counts = lambda r: [(s["newly_reachable"], s["newly_certified"]) for s in r.iteration_stats]
    assert counts(a) == counts(b), label


@pytest.mark.gpu
def test_static_run_matches_allocating_run_twice(gpu):
    first = cases.unicycle_30x30x24x10(sample_time=0.5)  # 216,000 states x 25 inputs
    second = cases.builtin("wraparound")                # 3,072 states x 4 inputs
    assert second.state_count() * second.input_count() < first.state_count() * first.input_count()
    for label, case in (("first, 30x30x24x10 x 25", first), ("second, wraparound 3072 x 4", second)):
        print(f"[case] no-alloc {label}", flush=True)
        want = run_case_u32(case, abstraction=True)
        got = run_case_u32_static(case, abstraction=True)
        _same(got, want, label, abstraction=True)


@pytest.mark.gpu
def test_static_run_without_export_and_size_limit(gpu):
    case = cases.builtin("default")
    print("[case] no-alloc builtin default without export", flush=True)
    _same(run_case_u32_static(case), run_case_u32(case), "default", abstraction=False)
    # 401M pairs cannot fit: rejected before any device work, with both sizes named.
    big = cases.unicycle_91x105x42x10()
    with pytest.raises(GscError) as e:
        run_case_u32_static(big)
    assert e.value.status == capi.GSC_ERR_OUT_OF_MEMORY
    assert "GSC_STATIC_SCRATCH_BYTES" in str(e.value) and "bytes of device memory" in str(e.value)
