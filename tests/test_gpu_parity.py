"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact for everything (integer sets, counts, tables and the FP64-derived
cell boxes): masks, abstraction triplets, reachable/candidate sets,
controller and reach_step tables, per-iteration counts, stop message.
"""
import numpy as np
import pytest

from casegen import assert_same_abstraction, assert_same_solution, overflow_regrowth_case, small_random_case
from gsc_amd import cases
from gsc_amd.solver import GscError, Solver, run_case_u32

pytestmark = pytest.mark.gpu


def _check(case, oracle, abstraction=True, label=None):
    label = label or case.name
    print(f"[case] {label} states={case.state_count()} inputs={case.input_count()}", flush=True)
    ref = oracle.run(case)
    got = run_case_u32(case, abstraction=abstraction)
    assert got.executed, got.message
    assert_same_solution(got, ref, label)
    if abstraction:
        assert_same_abstraction(got, ref, label)
    return got, ref


@pytest.mark.parametrize("which", ["default", "wraparound", "smoke"])
def test_builtin_cases(gpu, oracle, which):
    # tests/test_small_case.cpp:48-56 runs the default and wraparound cases CPU vs GPU.
    _check(cases.builtin(which), oracle)


def test_hyperbolic_builtin_first_iterations(gpu, oracle):
    c = cases.builtin("hyperbolic")
    c.max_iterations = 3
    _check(c, oracle)


@pytest.mark.parametrize("T", [0.1, 0.5])
def test_unicycle_30x30x24x10(gpu, oracle, T):
    _check(cases.unicycle_30x30x24x10(sample_time=T), oracle)


@pytest.mark.parametrize("seed", list(range(24)))
def test_small_random(gpu, oracle, seed):
    c = small_random_case(seed)
    try:
        ref_masks = oracle.masks(c)
    except ValueError as e:
        with pytest.raises(GscError) as ei:
            run_case_u32(c)
        assert str(ei.value) == str(e)
        return
    _check(c, oracle)


@pytest.mark.parametrize("n0,n1", [(100, 128), (100, 129), (200, 70)])
def test_plane_size_limit(gpu, oracle, n0, n1):
    # The solver builds the x/y prefix of a (θ, v) plane in LDS when the plane,
    # with rows padded to 64 cells, fits in 64 KiB (100x128 exactly fits), and
    # by separate row / column passes otherwise (100x129, 200x70).
    lb, ub, eta = cases._domain((n0, n1, 24, 3))
    ilb, iub, ieta = cases._inputs(3)
    c = cases.Case(
        name=f"plane_{n0}x{n1}", state_lb=lb, state_ub=ub, state_eta=eta, input_lb=ilb, input_ub=iub,
        input_eta=ieta, candidate=cases._r((4.0, 4.0, -cases.PI_LITERAL, 0.0), (6.0, 6.0, cases.PI_LITERAL, 2.0)),
        map=cases._r(lb, ub), obstacles=[cases._r((1.0, 7.0, lb[2], lb[3]), (3.0, 8.0, ub[2], ub[3]))],
        sample_time=0.5, integration_substeps=32, max_iterations=25, disturbance_half_width=(0.0, 0.0, 0.0, 0.0),
    )
    got, _ = _check(c, oracle)
    assert got.iterations == 25


def test_sweep_all_pairs_schedule(gpu, oracle):
    # The reference's schedule (every pair tested every iteration) must give the same result.
    c = cases.unicycle_30x30x24x10(sample_time=0.5)
    ref = oracle.run(c)
    with Solver(c, sweep_all_pairs=True) as s:
        s.build_abstraction()
        summ = s.solve()
        got = s.download()
    assert_same_solution(got, ref, "sweep")
    assert summ["pairs_tested"] == c.state_count() * c.input_count() * summ["iterations"]


RECOMPUTE_SEEDS = list(range(16))


@pytest.mark.parametrize("which", ["builtin:default", "builtin:wraparound", "builtin:smoke", "c1"]
                         + [f"seed:{s}" for s in RECOMPUTE_SEEDS])
def test_recompute_abstraction_mode(gpu, oracle, which):
    # No stored pair records: every tested box is recomputed in FP64 each iteration.
    if which.startswith("builtin:"):
        c = cases.builtin(which.split(":")[1])
        c.max_iterations = min(c.max_iterations, 6)
    elif which == "c1":
        c = cases.unicycle_30x30x24x10(sample_time=0.5)
    else:
        c = small_random_case(int(which.split(":")[1]))
        try:
            oracle.masks(c)
        except ValueError:
            pytest.skip("configuration error case")
    print(f"[case] recompute {which}", flush=True)
    ref = oracle.run(c, keep_abstraction=False)
    with Solver(c, recompute_abstraction=True) as s:
        assert s.build_abstraction() >= 0.0  # state hulls only
        s.solve()
        got = s.download()
        with pytest.raises(GscError, match="not stored in recompute mode"):
            s.download(abstraction=True)
    assert_same_solution(got, ref, f"recompute {which}")
    assert [st["newly_reachable"] for st in got.iteration_stats] == [int(v) for v in ref["newly_reachable"]]


def test_recompute_sweep_all_pairs(gpu, oracle):
    c = cases.unicycle_30x30x24x10(sample_time=0.5)
    ref = oracle.run(c, keep_abstraction=False)
    with Solver(c, recompute_abstraction=True, sweep_all_pairs=True) as s:
        summ = s.solve()
        got = s.download()
    assert_same_solution(got, ref, "recompute sweep")
    assert summ["pairs_tested"] == c.state_count() * c.input_count() * summ["iterations"]


def test_resolve_is_idempotent(gpu, oracle):
    # Re-seeding and re-solving on resident state gives the same answer twice.
    c = cases.unicycle_30x30x24x10(sample_time=0.5)
    ref = oracle.run(c)
    with Solver(c) as s:
        s.build_abstraction()
        s.solve()
        a = s.download()
        s.solve()
        b = s.download()
    assert_same_solution(a, ref, "first")
    assert_same_solution(b, ref, "second")


def test_candidate_outside_valid_raises(gpu):
    c = cases.builtin("default")
    # Cell centres sit at ±0.25: this obstacle removes four candidate columns.
    c.obstacles = [((-0.3, -0.3, -4.0, 0.0), (0.3, 0.3, 4.0, 2.0))]
    with pytest.raises(GscError, match="candidate set contains invalid states"):
        run_case_u32(c)


def test_max_iterations_zero(gpu, oracle):
    c = cases.unicycle_30x30x24x10(sample_time=0.5)
    c.max_iterations = 0
    ref = oracle.run(c)
    got = run_case_u32(c)
    assert_same_solution(got, ref, "max0")
    assert got.result_message == "maximum iterations reached"


def test_overflow_table_regrowth(gpu, oracle):
    """More escaped boxes than the first overflow table holds: the abstraction
    frees that table, allocates one of the needed size and runs again; the
    triplet (escaped boxes included) and the solve stay bit-exact."""
    case = overflow_regrowth_case()
    got, ref = _check(case, oracle, abstraction=True, label="overflow regrowth")
    mn, mx, va = ref["min_flat"], ref["max_flat"], ref["pair_valid"].astype(bool)
    n0 = 64  # x cells
    wrapped = va & ((mx % n0) < (mn % n0))
    assert wrapped.sum() > 65536  # the case really needs the regrowth
    assert got.memory_stats["abstraction_bytes"] >= case.state_count() * case.input_count() * 4 + wrapped.sum() * 16


# Frontier skipping has three modes (gsc_solve_summary::frontier): per-plane
# bitmasks where every box fits the bitmask row scans, prefix-table hull
# counts otherwise, and none under wrap on x or y. Each mode is pinned on
# cases that take it, and each result is checked against the oracle, so that
# a change that moves a case to another mode shows here as well as in timing.
_FRONTIER_CASES = [
    ("builtin:default", 2), ("builtin:wraparound", 2), ("c1", 2), ("small:2", 2), ("small:9", 2),
    ("small:3", 1), ("small:19", 1), ("small:5", 0), ("small:21", 0),
]


@pytest.mark.parametrize("which,mode", _FRONTIER_CASES, ids=[w for w, _ in _FRONTIER_CASES])
def test_frontier_mode(gpu, oracle, which, mode):
    kind, _, arg = which.partition(":")
    if kind == "builtin":
        case = cases.builtin(arg)
    elif kind == "c1":
        case = cases.unicycle_30x30x24x10(sample_time=0.5)
    else:
        case = small_random_case(int(arg))
    print(f"[case] {which} wrap_dim={case.wrap_dim}", flush=True)
    ref = oracle.run(case)
    with Solver(case, time_kernels=False) as s:
        s.build_abstraction()
        summary = s.solve()
        got = s.download(abstraction=False)
    assert summary["frontier"] == mode, f"{which}: frontier mode {summary['frontier']}, expected {mode}"
    assert_same_solution(got, ref, which)


# Every satisfaction tile (cells per workgroup, inputs per block) on cases that
# run each frontier mode: the bitmask frontier (builtin default, c1, small:2),
# the prefix-table frontier (small:3) and none (small:5, x/y wrap), and the
# sweep schedule on one of them.
_TILE_CASES = ["builtin:default", "c1", "small:2", "small:3", "small:5"]


@pytest.mark.parametrize("cells", [128, 256, 512, 1024])
@pytest.mark.parametrize("ib", [4, 8])
def test_satisfy_tiles(gpu, oracle, cells, ib):
    for which in _TILE_CASES:
        kind, _, arg = which.partition(":")
        if kind == "builtin":
            case = cases.builtin(arg)
        elif kind == "c1":
            case = cases.unicycle_30x30x24x10(sample_time=0.5)
        else:
            case = small_random_case(int(arg))
        print(f"[case] {which} tile {cells}x{ib}", flush=True)
        ref = oracle.run(case)
        for sweep in ((False, True) if which == "small:2" else (False,)):
            with Solver(case, time_kernels=False, sweep_all_pairs=sweep, tile_wg=cells, tile_ib=ib) as s:
                s.build_abstraction()
                s.solve()
                got = s.download(abstraction=False)
            assert_same_solution(got, ref, f"{which} tile {cells}x{ib} sweep={sweep}")


def test_satisfy_tile_rejected(gpu):
    with pytest.raises(GscError):
        Solver(cases.builtin("default"), tile_wg=64)
    with pytest.raises(GscError):
        Solver(cases.builtin("default"), tile_ib=16)
