"""Serving with continuous batching (gsc_server_*): every result equals the
oracle and is the same alone as in a mix of concurrent requests; admission
respects the HBM budget and the slot count; failures stay per request."""
import numpy as np
import pytest

from casegen import SOLVE_KEYS, assert_same_solution, uneven_cases
from gsc_amd import cases
from gsc_amd.serve import Server, poisson_load

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mix(gpu, oracle):
    cs = uneven_cases(oracle)
    refs = [oracle.run(c, keep_abstraction=False) for c in cs]
    return cs, refs


def _same(a, b, label):
    for k in SOLVE_KEYS:
        assert np.array_equal(getattr(a, k), getattr(b, k)), f"{label}: {k}"
    assert a.iterations == b.iterations and a.result_message == b.result_message, label
    assert [s["newly_reachable"] for s in a.iteration_stats] == [s["newly_reachable"] for s in b.iteration_stats]


def test_mix_equals_alone_equals_oracle(mix):
    cs, refs = mix
    with Server(max_active=5, iterations_per_step=3, hbm_budget_bytes=4 << 30) as srv:
        print(f"[case] serve mix of {len(cs)}", flush=True)
        ids = [srv.submit(c) for c in cs]  # all at once: queueing, staggered admission, joins mid-run
        in_mix = []
        for c, ref, rid in zip(cs, refs, ids):
            info = srv.wait(rid)
            assert info["state"] == 2, info["error"]
            r = srv.fetch(rid)
            assert_same_solution(r, ref, f"mix {c.name}")
            assert [s["newly_reachable"] for s in r.iteration_stats] == [int(v) for v in ref["newly_reachable"]]
            in_mix.append(r)
        st = srv.stats()
        assert st["completed"] == len(cs) and st["failed"] == 0
        assert 1 < st["max_active_seen"] <= 5
        assert st["budget_in_use"] == 0  # every case returned its HBM
        print("[case] serve alone", flush=True)
        for c, r_mix in zip(cs, in_mix):
            rid = srv.submit(c)
            assert srv.wait(rid)["state"] == 2
            _same(srv.fetch(rid), r_mix, f"alone {c.name}")


def test_budget_limits_admission(mix):
    cs, refs = mix
    big = cases.unicycle_30x30x24x10(sample_time=0.5)
    with Server(max_active=64, iterations_per_step=2, hbm_budget_bytes=64 << 20) as srv:
        print("[case] serve small budget", flush=True)
        ids = [srv.submit(big) for _ in range(6)]
        infos = [srv.wait(r) for r in ids]
        assert all(i["state"] == 2 for i in infos), [i["error"] for i in infos]
        per_case = infos[0]["device_bytes"]
        st = srv.stats()
        assert st["budget_peak"] <= st["budget_bytes"]
        assert st["max_active_seen"] <= max(1, st["budget_bytes"] // per_case)
        ref = refs[cs.index(next(c for c in cs if c.name == big.name))]
        for rid in ids:
            assert_same_solution(srv.fetch(rid), ref, "budget")
        # A case larger than the whole budget fails on its own; the server keeps serving.
        huge = cases.unicycle_91x105x42x10()
        bad = srv.submit(huge)
        info = srv.wait(bad)
        assert info["state"] == 3 and "budget" in info["error"]
        ok = srv.submit(big)
        assert srv.wait(ok)["state"] == 2
        assert_same_solution(srv.fetch(ok), ref, "after failure")


def test_config_error_fails_only_that_request(mix):
    cs, refs = mix
    bad = cases.unicycle_30x30x24x10(sample_time=0.5)
    bad.obstacles = [((4.5, 4.5, -4.0, -1.0), (5.5, 5.5, 4.0, 3.0))]
    with Server(max_active=8) as srv:
        print("[case] serve config error", flush=True)
        rb = srv.submit(bad)
        rg = srv.submit(cs[0])
        ib = srv.wait(rb)
        assert ib["state"] == 3 and "candidate set contains invalid states" in ib["error"]
        assert srv.wait(rg)["state"] == 2
        assert_same_solution(srv.fetch(rg), refs[0], "neighbour of a failed request")


def test_poisson_stream(mix):
    cs, refs = mix
    with Server(max_active=16, iterations_per_step=4, hbm_budget_bytes=8 << 30) as srv:
        print("[case] serve poisson", flush=True)
        stats, results = poisson_load(srv, cs, rate_per_s=400.0, seed=7, keep_results=True)
        assert not stats["failed"] and stats["requests"] == len(cs)
        assert stats["p50_ms"] <= stats["p99_ms"]
        for c, ref, r in zip(cs, refs, results):
            assert_same_solution(r, ref, f"poisson {c.name}")


def test_shutdown_with_queued_work(mix):
    cs, _ = mix
    srv = Server(max_active=1, iterations_per_step=1)
    ids = [srv.submit(c) for c in cs]
    srv.close()  # running case finishes, queued ones fail with a shutdown error; no hang
    assert srv.handle is None
