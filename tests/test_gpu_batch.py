"""Batched solves: many uneven cases advanced together (gsc_batch_*), each
identical to the oracle and to its own single-case solve, under every tile and
schedule configuration the autotuner may pick, on repeated graph replays."""
import numpy as np
import pytest

from casegen import assert_same_solution, uneven_cases
from gsc_amd import cases
from gsc_amd.batch import Batch
from gsc_amd.solver import GscError

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def batch_and_refs(gpu, oracle):
    cs = uneven_cases(oracle)
    refs = [oracle.run(c, keep_abstraction=False) for c in cs]
    b = Batch(cs)
    b.build_abstraction()
    yield b, cs, refs
    b.close()


def _check_all(b, cs, refs, label):
    for k, (c, ref) in enumerate(zip(cs, refs)):
        got = b.download(k)
        assert got.executed, got.message
        assert_same_solution(got, ref, f"{label} {c.name}")
        newly = [s["newly_reachable"] for s in got.iteration_stats]
        assert newly == [int(v) for v in ref["newly_reachable"]], f"{label} {c.name}"
        cert = [s["newly_certified"] for s in got.iteration_stats]
        assert cert == [int(v) for v in ref["newly_certified"]], f"{label} {c.name}"
        s = b.case_summary(k)
        assert s["iterations"] == ref["iterations"] and s["message"] == ref["message"]
        assert s["reachable_states"] == int(ref["reachable_mask"].sum())


def test_batch_matches_oracle(batch_and_refs):
    b, cs, refs = batch_and_refs
    print(f"[case] batch of {len(cs)} default options", flush=True)
    summ = b.solve()
    assert summ["cases"] == len(cs) and summ["stopped"] == len(cs)
    assert summ["max_iterations"] == max(r["iterations"] for r in refs)
    _check_all(b, cs, refs, "default")


@pytest.mark.parametrize("wg", [128, 256, 512, 1024])
@pytest.mark.parametrize("ib", [4, 8])
def test_batch_tile_variants(batch_and_refs, wg, ib):
    b, cs, refs = batch_and_refs
    print(f"[case] batch tile wg={wg} ib={ib}", flush=True)
    b.set_options(tile_wg=wg, tile_ib=ib)
    b.solve()
    _check_all(b, cs, refs, f"wg{wg} ib{ib}")


@pytest.mark.parametrize("g,order,graph", [(1, 0, 1), (3, 1, 1), (32, 0, 1), (8, 1, 0)])
def test_batch_schedules(batch_and_refs, g, order, graph):
    b, cs, refs = batch_and_refs
    print(f"[case] batch schedule G={g} order={order} graph={graph}", flush=True)
    b.set_options(tile_wg=256, tile_ib=8, iterations_per_launch=g, order=order, use_graph=graph)
    for rep in range(2):  # the captured graph is replayed across solves
        b.solve()
        _check_all(b, cs, refs, f"G{g} order{order} graph{graph} rep{rep}")


def test_batch_tune_keeps_results(batch_and_refs):
    b, cs, refs = batch_and_refs
    print("[case] batch tune", flush=True)
    tried = b.tune()
    assert len(tried) >= 10
    best = min(tried, key=lambda e: e["solve_ms"])
    opts = b.options
    assert (opts["tile_wg"], opts["tile_ib"], opts["iterations_per_launch"], opts["order"]) == (
        best["tile_wg"], best["tile_ib"], best["iterations_per_launch"], best["order"])
    _check_all(b, cs, refs, "tuned")
    b.solve()
    _check_all(b, cs, refs, "tuned again")


def test_batch_config_error_names_case(gpu):
    good = cases.unicycle_30x30x24x10(sample_time=0.5)
    bad = cases.unicycle_30x30x24x10(sample_time=0.5)
    bad.obstacles = [((4.5, 4.5, -4.0, -1.0), (5.5, 5.5, 4.0, 3.0))]  # inside the target: candidate not valid
    with pytest.raises(GscError) as ei:
        Batch([good, bad])
    assert "case 1" in str(ei.value) and "candidate set contains invalid states" in str(ei.value)


def _wide_plane_case():
    # 100 x 129 cells per (θ, v) plane: 128-cell rows x 129 rows of 4 bytes
    # exceed the 64 KiB of LDS the fused prefix build uses.
    lb, ub, eta = cases._domain((100, 129, 24, 3))
    ilb, iub, ieta = cases._inputs(3)
    return cases.Case(
        name="plane_100x129", state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub, input_eta=ieta,
        candidate=cases._r((4.0, 4.0, -cases.PI_LITERAL, 0.0), (6.0, 6.0, cases.PI_LITERAL, 2.0)),
        map=cases._r(lb, ub), obstacles=[cases._r((1.0, 7.0, lb[2], lb[3]), (3.0, 8.0, ub[2], ub[3]))],
        sample_time=0.5, integration_substeps=32, max_iterations=25, disturbance_half_width=(0.0, 0.0, 0.0, 0.0),
    )


def test_batch_with_a_plane_too_large_for_lds(gpu, oracle):
    # One case whose plane does not fit in LDS sends the whole batch through the
    # separate prefix passes; every case must still match the oracle.
    cs = [cases.unicycle_30x30x24x10(sample_time=0.5), _wide_plane_case(), cases.builtin("default")]
    print(f"[case] batch with a wide plane: {[c.name for c in cs]}", flush=True)
    refs = [oracle.run(c, keep_abstraction=False) for c in cs]
    with Batch(cs) as b:
        b.build_abstraction()
        b.solve()
        _check_all(b, cs, refs, "wide plane")
