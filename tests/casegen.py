"""Small seeded cases for parity tests (both the CPU oracle suite and the GPU suite).

They stress the corners of the reference's arithmetic: every wrap_dim, both
state constraints, large disturbances (wrapped θ boxes, full-range θ boxes),
odd substep counts, extension targets, and boxes that leave the domain.
"""
import math

import numpy as np

from gsc_amd import capi
from gsc_amd.cases import PI_LITERAL, Case, _r, fit_eta


def small_random_case(seed: int, max_iterations: int = 60) -> Case:
    rng = np.random.default_rng(1000 + seed)
    shape = [int(rng.integers(3, 13)) for _ in range(4)]
    lb = [float(rng.uniform(-3, 0)), float(rng.uniform(-3, 0)), -PI_LITERAL, float(rng.uniform(-0.5, 0.5))]
    ub = [lb[0] + float(rng.uniform(2, 6)), lb[1] + float(rng.uniform(2, 6)), PI_LITERAL, lb[3] + float(rng.uniform(1, 3))]
    # Some cases use a plain (possibly inexact) eta, as hand-written configs do.
    if rng.random() < 0.5:
        eta = [fit_eta(ub[d] - lb[d], shape[d]) for d in range(4)]
    else:
        eta = [round((ub[d] - lb[d]) / shape[d], 3) for d in range(4)]
    wrap_dim = int(rng.choice([2, 2, 2, 0, 1, 3]))
    ni = [int(rng.integers(1, 6)), int(rng.integers(1, 6))]
    ilb = [float(rng.uniform(-1.5, 0)), float(rng.uniform(-1.5, 0))]
    iub = [ilb[0] + float(rng.uniform(0.5, 3)), ilb[1] + float(rng.uniform(0.5, 3))]
    ieta = [(iub[d] - ilb[d]) / ni[d] for d in range(2)]
    T = float(rng.choice([0.05, 0.1, 0.3, 0.5, 1.0]))
    dw = [float(rng.uniform(0, 0.1)), float(rng.uniform(0, 0.1)), float(rng.choice([0.0, 0.02, 0.35, 1.5])),
          float(rng.uniform(0, 0.05))]
    cx, cy = (lb[0] + ub[0]) / 2, (lb[1] + ub[1]) / 2
    hw = float(rng.uniform(0.3, 1.2))
    cand = _r((cx - hw, cy - hw, -PI_LITERAL, lb[3]), (cx + hw, cy + hw, PI_LITERAL, ub[3]))
    obstacles = []
    for _ in range(int(rng.integers(0, 3))):
        ox = float(rng.choice([lb[0] + 0.2, ub[0] - 1.0]))
        oy = float(rng.uniform(lb[1], ub[1] - 0.5))
        obstacles.append(_r((ox, oy, -PI_LITERAL, lb[3]), (ox + 0.6, oy + 0.5, PI_LITERAL, ub[3])))
    constraint = int(rng.choice([0, 0, 1, 2]))
    case = Case(
        name=f"small_{seed}", state_lb=tuple(lb), state_ub=tuple(ub), state_eta=tuple(eta), wrap_dim=wrap_dim,
        input_lb=tuple(ilb), input_ub=tuple(iub), input_eta=tuple(ieta), candidate=cand, map=_r(lb, ub),
        obstacles=obstacles, constraint_type=constraint,
        hyperbolic_a=float(rng.uniform(2, 6)), hyperbolic_b=float(rng.uniform(1, 4)),
        hyperbolic_c=float(rng.uniform(4, 20)), elliptic_a=float(rng.uniform(1.5, 4)),
        elliptic_b=float(rng.uniform(1.5, 4)), disturbance_half_width=tuple(dw), sample_time=T,
        integration_substeps=int(rng.choice([1, 7, 32])), max_iterations=max_iterations, verbose=False,
    )
    if rng.random() < 0.3:
        case.extra_targets = [_r((lb[0], lb[1], -PI_LITERAL, lb[3]), (lb[0] + 0.7, lb[1] + 0.7, PI_LITERAL, ub[3]))]
    if rng.random() < 0.2:
        case.target_band = (cx * cy, 0.5)
    # Keep the candidate set inside the valid set where the constraint allows it.
    if constraint == capi.CONSTRAINT_HYPERBOLIC:
        case.hyperbolic_a = max(case.hyperbolic_a, (abs(cx) + hw + 1) ** 2)
        case.hyperbolic_c = max(case.hyperbolic_c, case.hyperbolic_b * (abs(cy) + hw + 1) ** 2)
    elif constraint == capi.CONSTRAINT_ELLIPTIC:
        r = math.hypot(abs(cx) + hw + 1, abs(cy) + hw + 1)
        case.elliptic_a = max(case.elliptic_a, 1.5 * r)
        case.elliptic_b = max(case.elliptic_b, 1.5 * r)
    return case


def overflow_regrowth_case() -> Case:
    """x wraps (wrap_dim 0) and the x disturbance is wide, so 212,480 of the
    326,656 valid successor boxes wrap in x and cannot be packed in a record
    word: more than three times the 65,536 overflow entries a solver starts
    with, so the abstraction regrows its overflow table and runs again."""
    lb = (0.0, 0.0, -PI_LITERAL, 0.0)
    ub = (6.4, 6.4, PI_LITERAL, 1.0)
    shape = (64, 64, 8, 4)
    eta = tuple(fit_eta(ub[d] - lb[d], shape[d]) for d in range(4))
    return Case(name="overflow_regrowth", state_lb=lb, state_ub=ub, state_eta=eta, wrap_dim=0,
                input_lb=(-1.0, -1.0), input_ub=(1.0, 1.0), input_eta=(1.0, 1.0),
                candidate=_r((0.0, 1.0, -PI_LITERAL, 0.0), (6.4, 5.4, PI_LITERAL, 1.0)), map=_r(lb, ub),
                disturbance_half_width=(3.0, 0.05, 0.02, 0.02), sample_time=0.5, max_iterations=40)


SOLVE_KEYS = ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step")


def assert_same_solution(gpu, ref, label=""):
    """Byte-identical arrays and identical iteration history (main.cpp:114-121 plus counts)."""
    for k in SOLVE_KEYS:
        a = getattr(gpu, k)
        b = ref[k]
        assert a.shape == b.shape, f"{label}: {k} shape {a.shape} vs {b.shape}"
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            raise AssertionError(f"{label}: {k} differs at {len(bad)} states, first {bad[:5]} gpu={a[bad[:5]]} ref={b[bad[:5]]}")
    assert gpu.iterations == ref["iterations"], f"{label}: iterations {gpu.iterations} vs {ref['iterations']}"
    assert gpu.converged == ref["converged"], label
    assert gpu.result_message == ref["message"], f"{label}: {gpu.result_message!r} vs {ref['message']!r}"
    nr = [s["newly_reachable"] for s in gpu.iteration_stats]
    nc = [s["newly_certified"] for s in gpu.iteration_stats]
    assert nr == [int(v) for v in ref["newly_reachable"]], f"{label}: newly_reachable {nr} vs {list(ref['newly_reachable'])}"
    assert nc == [int(v) for v in ref["newly_certified"]], f"{label}: newly_certified {nc} vs {list(ref['newly_certified'])}"
    assert gpu.reachable_states == int(ref["reachable_mask"].sum()), label
    assert gpu.certified_candidate_states == int(np.sum(ref["newly_certified"])), label


def assert_same_abstraction(gpu, ref, label=""):
    for k_gpu, k_ref in (("abstraction_valid", "pair_valid"), ("abstraction_min_flat", "min_flat"),
                         ("abstraction_max_flat", "max_flat")):
        a = getattr(gpu, k_gpu)
        b = ref[k_ref]
        if not np.array_equal(a, b):
            bad = np.flatnonzero(a != b)
            raise AssertionError(f"{label}: {k_gpu} differs at {len(bad)} pairs, first {bad[:5]} gpu={a[bad[:5]]} ref={b[bad[:5]]}")


def uneven_cases(oracle, n_small: int = 8, n_random: int = 4):
    """Cases of uneven grid and input sizes for the batched and served paths:
    valid small random cases (configuration errors skipped), 16..100-input
    instances with ~9:10:4:1 grids, the 30x30x24x10 case, a case whose
    overflow table regrows (overflow_regrowth_case) and a case with
    max_iterations = 0."""
    from gsc_amd import cases as gcases

    out = []
    for seed in range(60):
        c = small_random_case(seed)
        try:
            oracle.masks(c)
        except ValueError:
            continue
        out.append(c)
        if len(out) == n_small:
            break
    rng = np.random.default_rng(11)
    for k in range(n_random):
        out.append(gcases.random_instance(rng, 3e3, 2e4, f"uneven_{k}"))
    out.append(gcases.unicycle_30x30x24x10(sample_time=0.5))
    out.append(overflow_regrowth_case())
    z = small_random_case(3)
    z.name = "zero_iterations"
    z.max_iterations = 0
    out.append(z)
    return out
