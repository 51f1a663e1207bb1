"""Pins the CPU oracle (the checker of every GPU parity test).

1. Against the reference's own artefacts for its published run
   (README.md:44-98 per-iteration counts; plots/reachable_set.mat digest).
2. Against golden fixtures written by the reference's own CPU build
   (tests/golden/make_goldens.py -> oracle/_ref), including the reference's
   canned cases built by its own make_*_case().
3. Live against the reference build on fresh random cases, where it is built.
4. Against closed forms: brute-force prefix sums and box counts, the
   ceil-shape rule, and invariants of the fixpoint.
"""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from casegen import small_random_case
from conftest import GOLDEN
from gsc_amd import cases


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_golden(name):
    with open(os.path.join(GOLDEN, f"{name}.json")) as f:
        return json.load(f)


def check_against(fx, out, label):
    for k, digest in fx["sha256"].items():
        assert sha(out[k]) == digest, f"{label}: {k} digest differs"
    assert out["iterations"] == fx["iterations"], label
    assert out["message"] == fx["message"], label
    assert [int(v) for v in out["newly_reachable"]] == fx["newly_reachable"], label
    assert [int(v) for v in out["newly_certified"]] == fx["newly_certified"], label


@pytest.mark.parametrize("which", ["default", "wraparound", "smoke"])
def test_builtin_cases_match_reference_make_case(oracle, which):
    """gsc_config_builtin + the oracle reproduce the reference's own canned case."""
    fx = load_golden(f"builtin_{which}")
    out = oracle.run(cases.builtin(which))
    check_against(fx, out, which)


def test_tiny_case_full_arrays(oracle):
    data = np.load(os.path.join(GOLDEN, "builtin_default.npz"))
    out = oracle.run(cases.builtin("default"))
    for k in data.files:
        assert np.array_equal(out[k], data[k]), k


@pytest.mark.parametrize("T", ["0.1", "0.5"])
def test_unicycle_30x30x24x10(oracle, T):
    fx = load_golden(f"unicycle_30x30x24x10_T{T}")
    out = oracle.run(cases.unicycle_30x30x24x10(sample_time=float(T)))
    check_against(fx, out, f"T={T}")


def test_small_random_cases(oracle):
    golden = load_golden("small_random")
    for seed, fx in golden.items():
        case = small_random_case(int(seed))
        if "error" in fx:
            with pytest.raises(ValueError) as ei:
                oracle.masks(case)
            assert str(ei.value) == fx["error"], seed
            continue
        check_against(fx, oracle.run(case), f"seed {seed}")


def test_published_case(oracle):
    """The reference's published 70x70x63x13 x 100 run (README.md:44-98)."""
    fx = load_golden("hyperbolic_demo")
    out = oracle.run(cases.hyperbolic_demo(), threads=min(8, os.cpu_count() or 1))
    check_against(fx, out, "hyperbolic_demo")
    # README.md:48-52 and the per-iteration lines :88-98 (iterations 6-7 elided there).
    assert out["reachable_states"] == 1142689
    assert int(out["candidate_mask"].sum()) == 161280
    assert out["iterations"] == 12 and out["converged"]
    readme_nr = {1: 52481, 2: 57431, 3: 91549, 4: 122322, 5: 129220, 8: 94948, 9: 78464, 10: 64643, 11: 43185,
                 12: 16768}
    readme_nc = {1: 123510, 2: 1097, 3: 1502, 4: 1501, 5: 1459, 8: 5579, 9: 7315, 10: 7976, 11: 3937, 12: 130}
    for it, v in readme_nr.items():
        assert int(out["newly_reachable"][it - 1]) == v
    for it, v in readme_nc.items():
        assert int(out["newly_certified"][it - 1]) == v
    # plots/reachable_set.mat: reachable_mask_3d, produced by the reference's
    # post-processing as any() over the last axis of a C-order (70,70,63,13)
    # reshape of the flat reachable mask.
    m = fx["mat_reachable_mask_3d"]
    mat = np.ascontiguousarray(out["reachable_mask"].reshape(70, 70, 63, 13).any(axis=3)).astype(np.uint8)
    assert list(mat.shape) == m["shape"]
    assert int(mat.sum()) == m["sum"]
    assert sha(mat) == m["sha256_C_order"]


def test_live_against_reference_build(oracle):
    if not oracle.ref_available():
        pytest.skip("reference CPU build not present (oracle/_ref is built only where the reference exists)")
    for seed in range(100, 106):
        case = small_random_case(seed)
        try:
            ref = oracle.run_ref(case)
        except RuntimeError as e:
            with pytest.raises(ValueError):
                oracle.masks(case)
            continue
        out = oracle.run(case)
        for k in ("valid_mask", "candidate_mask", "reachable_mask", "controller", "reach_step", "min_flat", "max_flat",
                  "pair_valid"):
            assert np.array_equal(out[k], ref[k]), (seed, k)
        assert out["iterations"] == ref["iterations"] and out["message"] == ref["message"], seed
        assert list(out["newly_reachable"]) == list(ref["newly_reachable"]), seed
        assert list(out["newly_certified"]) == list(ref["newly_certified"]), seed


# ---- closed forms ------------------------------------------------------------
def _tiny_case(shape):
    c = cases.builtin("default")
    c.state_lb = (0.0, 0.0, 0.0, 0.0)
    c.state_ub = tuple(float(n) for n in shape)
    c.state_eta = (1.0, 1.0, 1.0, 1.0)
    return c


def test_prefix_sum_is_4d_cumsum(oracle):
    rng = np.random.default_rng(7)
    for shape in [(3, 4, 5, 2), (7, 1, 3, 6), (1, 1, 1, 1), (5, 5, 5, 5)]:
        c = _tiny_case(shape)
        mask = (rng.random(int(np.prod(shape))) < 0.4).astype(np.uint8)
        table = oracle.prefix_sum(c, mask)
        m4 = mask.reshape(shape[::-1])  # [w][z][y][x], x fastest
        cs = m4.astype(np.uint64).cumsum(0).cumsum(1).cumsum(2).cumsum(3)
        padded = np.zeros(tuple(n + 1 for n in shape[::-1]), np.uint64)
        padded[1:, 1:, 1:, 1:] = cs
        assert np.array_equal(table, padded.reshape(-1)), shape


def test_box_count_matches_brute_force(oracle):
    rng = np.random.default_rng(11)
    shape = (6, 5, 7, 4)
    c = _tiny_case(shape)
    c.wrap_dim = 2
    S = int(np.prod(shape))
    mask = (rng.random(S) < 0.5).astype(np.uint8)
    table = oracle.prefix_sum(c, mask)
    m4 = mask.reshape(shape[::-1])
    mins, maxs, want = [], [], []
    strides = (1, shape[0], shape[0] * shape[1], shape[0] * shape[1] * shape[2])
    for _ in range(400):
        lo = [int(rng.integers(0, n)) for n in shape]
        hi = [int(rng.integers(l, n)) for l, n in zip(lo, shape)]
        if rng.random() < 0.4:  # wrapped θ range: min > max
            lo[2], hi[2] = max(lo[2], hi[2]), min(lo[2], hi[2])
            if lo[2] == hi[2]:
                continue
        zs = list(range(lo[2], hi[2] + 1)) if lo[2] <= hi[2] else list(range(lo[2], shape[2])) + list(range(0, hi[2] + 1))
        cnt = int(m4[lo[3]:hi[3] + 1, zs][:, :, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].sum())
        mins.append(sum(l * s for l, s in zip(lo, strides)))
        maxs.append(sum(h * s for h, s in zip(hi, strides)))
        want.append(cnt)
    got = oracle.query_box_count(c, table, np.array(mins), np.array(maxs))
    assert [int(v) for v in got] == want


def test_shape_rule_and_ceil_trap(oracle):
    # grid.hpp:121: ceil((ub - lb) / eta); (1.1 - (-0.1)) / 0.1 = 12.000000000000002 -> 13 cells.
    sh = oracle.shape(cases.hyperbolic_demo())
    assert sh["state_shape"] == (70, 70, 63, 13)
    assert sh["pair_count"] == 401_310_000
    assert sh["prefix_count"] == 71 * 71 * 64 * 14
    assert math.ceil((1.1 - (-0.1)) / 0.1) == 13
    for fn, want in ((cases.unicycle_30x30x24x10, (30, 30, 24, 10)),
                     (cases.unicycle_91x105x42x10, (91, 105, 42, 10)),
                     (cases.unicycle_128x128x64x16, (128, 128, 64, 16))):
        assert oracle.shape(fn())["state_shape"] == want


def test_fixpoint_invariants(oracle):
    """Properties fixed by the algorithm (reachability.hpp:30-155), independent of the reference run."""
    c = cases.unicycle_30x30x24x10(sample_time=0.5)
    out = oracle.run(c)
    valid, cand, reach = out["valid_mask"], out["candidate_mask"], out["reachable_mask"]
    ctrl, rstep = out["controller"], out["reach_step"]
    I = c.input_count()
    assert np.all(reach[cand == 1] == 1)                      # candidates seed the reachable set
    assert np.all(valid[reach == 1] == 1)                     # reachable ⊆ valid
    noncand = (reach == 1) & (cand == 0)
    assert np.all(ctrl[noncand] < I)                          # every reached state has a controller
    assert np.all((rstep[noncand] >= 1) & (rstep[noncand] <= out["iterations"]))
    assert np.all(rstep[cand == 1] == 0)
    assert np.all(rstep[reach == 0] == 0xFFFFFFFF)
    hist = np.bincount(rstep[noncand], minlength=out["iterations"] + 1)[1:]
    assert [int(v) for v in hist] == [int(v) for v in out["newly_reachable"]]
    # The chosen controller is the lowest satisfying input: re-check each reached
    # state's successor box lies inside the final reachable set (over valid cells).
    mn, mx, va = out["min_flat"], out["max_flat"], out["pair_valid"]
    pv = oracle.prefix_sum(c, valid)
    pr = oracle.prefix_sum(c, reach)
    idx = np.flatnonzero(noncand)[:300]
    pairs = idx * I + ctrl[idx]
    assert np.all(va[pairs] == 1)
    vc = oracle.query_box_count(c, pv, mn[pairs], mx[pairs])
    rc = oracle.query_box_count(c, pr, mn[pairs], mx[pairs])
    assert np.all(vc > 0) and np.all(rc == vc)
