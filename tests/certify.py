"""Size-independent certificate of a fixpoint result, against the oracle's own
successor boxes and box queries (reachability.hpp:30-155 semantics).

For a sample of states (contiguous chunks at seeded positions) the oracle
computes every pair's successor box (abstraction.hpp:161-209) and the check
replays the reference's decision for that state from the result alone:

  R_j = candidates  +  non-candidates with reach_step <= j     (Jacobi sets)
  sat(s, u, R) = valid pair  and  count_V(box) > 0  and  count_R(box) == count_V(box)

  - a non-candidate reached at iteration k: its controller u is the LOWEST
    input with sat(s, u, R_{k-1}), and no input is sat w.r.t. R_{k-2}
    (otherwise it would have been reached one iteration earlier);
  - a certified candidate: its controller is sat w.r.t. the final set;
  - an unreached non-candidate, or an uncertified candidate: no input is sat
    w.r.t. R_{N-1}, the set the last iteration used;
plus the set identities (reach within valid, candidates reached, reach_step 0
on candidates and unreachable elsewhere). Nothing here restates the GPU
kernels: boxes and counts come from the oracle."""
import numpy as np

INVALID = 0xFFFFFFFF


def _jacobi_set(cand, reach, rstep, j):
    return (cand != 0) | ((reach != 0) & (cand == 0) & (rstep <= j))


def sample_starts(S, n_chunks=40, chunk=64, seed=0):
    rng = np.random.default_rng(seed)
    return np.unique(rng.integers(0, max(1, S - chunk), size=n_chunks))


def certify(oracle, case, res, n_chunks=40, chunk=64, seed=0):
    valid = np.asarray(res.valid_mask)
    cand = np.asarray(res.candidate_mask)
    reach = np.asarray(res.reachable_mask)
    ctrl = np.asarray(res.controller)
    rstep = np.asarray(res.reach_step)
    N = int(res.iterations)
    S = valid.size
    I = case.input_count()
    # Set identities.
    assert not np.any(reach & ~valid.astype(bool)), "reached state outside the valid set"
    assert np.all(reach[cand != 0] == 1), "candidate not in the reachable set"
    assert np.all(rstep[cand != 0] == 0), "candidate reach_step must be 0"
    nc_reached = (reach != 0) & (cand == 0)
    assert np.all(rstep[(reach == 0)] == INVALID), "unreached state with a reach step"
    assert np.all((rstep[nc_reached] >= 1) & (rstep[nc_reached] <= max(N, 0))), "reach_step outside 1..N"
    assert np.all(ctrl[nc_reached] != INVALID), "reached state without a controller"
    # Sampled states.
    starts = sample_starts(S, n_chunks, chunk, seed)
    states, boxes = [], []
    for b in starts:
        e = min(S, int(b) + chunk)
        mn, mx, va = oracle.abstraction(case, int(b), e, threads=8)
        for k, s in enumerate(range(int(b), e)):
            states.append(s)
            boxes.append((mn[k * I:(k + 1) * I], mx[k * I:(k + 1) * I], va[k * I:(k + 1) * I]))
    pv = oracle.prefix_sum(case, valid)
    cnt_v = [oracle.query_box_count(case, pv, mn, mx) for mn, mx, _ in boxes]
    need = {}  # iteration index j -> list of sampled positions needing R_j
    for p, s in enumerate(states):
        if not valid[s]:
            continue
        if nc_reached[s]:
            k = int(rstep[s])
            need.setdefault(k - 1, []).append(p)
            if k >= 2:
                need.setdefault(k - 2, []).append(p)
        elif cand[s] and ctrl[s] != INVALID:
            need.setdefault(N, []).append(p)
        elif N >= 1:
            need.setdefault(N - 1, []).append(p)
    checked = 0
    for j in sorted(need):
        R = _jacobi_set(cand, reach, rstep, j).astype(np.uint8)
        pr = oracle.prefix_sum(case, R)
        for p in need[j]:
            s = states[p]
            mn, mx, va = boxes[p]
            cr = oracle.query_box_count(case, pr, mn, mx)
            sat = (va != 0) & (cnt_v[p] > 0) & (cr == cnt_v[p])
            first = int(np.argmax(sat)) if sat.any() else INVALID
            if nc_reached[s]:
                k = int(rstep[s])
                if j == k - 1:
                    assert first == int(ctrl[s]), f"state {s}: controller {ctrl[s]}, lowest satisfied input {first}"
                else:
                    assert first == INVALID, f"state {s} reached at {k} but input {first} satisfied at {k - 1}"
            elif cand[s] and ctrl[s] != INVALID:
                assert sat[int(ctrl[s])], f"candidate {s}: controller {ctrl[s]} not satisfied by the final set"
            else:
                assert first == INVALID, f"state {s} left unreached but input {first} satisfied at iteration {N}"
            checked += 1
    return {"sampled_states": len(states), "checked": checked, "sets": len(need)}
