"""GPU: the continuous-batching server.  Every request's records equal the
oracle's run of the same program bit for bit, and equal its records when run
alone (solo vs mix), whatever path (resident / stepwise), tick length or
admission order the scheduler takes."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sv():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd import serving

    return serving


def oracle_records(req):
    if req.kind == "wealth":
        x, y, w, rng = O.wealth_init(req.n, req.seed, O.MODE_PARTITIONABLE)
        mean, gini = O.wealth_run(x, y, w, rng, req.steps, O.MODE_PARTITIONABLE)
        return {"mean": mean, "gini": gini}
    rp, col = O.ba_csr(req.n, req.m, req.graph_seed)
    state, rng = O.sir_init(req.n, req.n_infected, req.seed, O.MODE_PARTITIONABLE)
    c = O.sir_run(rp, col, state, rng, np.float32(req.beta), np.float32(req.gamma), req.steps,
                  O.MODE_PARTITIONABLE)
    return {"S": c[:, 0], "I": c[:, 1], "R": c[:, 2]}


def same(a, b):
    return a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a)


def requests(sv):
    R = sv.SimRequest
    return [
        R("wealth", n=1000, steps=40, seed=0),
        R("wealth", n=3001, steps=25, seed=1),
        R("sir", n=2000, steps=60, seed=2, graph_seed=7, beta=0.08, gamma=0.1),
        R("sir", n=5003, steps=33, seed=3, graph_seed=8, beta=0.05, gamma=0.07, n_infected=20),
        R("wealth", n=2, steps=9, seed=4),
        R("sir", n=700, steps=20, seed=5, graph_seed=9, m=3, beta=0.2, gamma=0.3),
        # a stepwise wealth step t is applied at sub-step t + 1 and recorded at
        # t + 2: runs whose last record falls at, before and after the end of
        # a 7-step tick, and a one-step run
        R("wealth", n=1500, steps=1, seed=6),
        R("wealth", n=1501, steps=5, seed=7),
        R("wealth", n=1502, steps=6, seed=8),
        R("wealth", n=1503, steps=12, seed=9),
    ]


@pytest.mark.parametrize("path", ["resident", "stepwise"])
def test_server_matches_oracle(sv, path):
    thr = 65536 if path == "resident" else 0
    cfg = dict(max_active=4, sir_resident_max=thr, wealth_resident_max=thr, steps_per_tick=7,
               sir_chunk=2048, wealth_chunk=512, hbm_budget=256 << 20)
    reqs = requests(sv)
    with sv.SimulationServer(cfg) as s:
        tickets = [s.submit(r) for r in reqs]
        s.drain(120)
        for t, r in zip(tickets, reqs):
            info = s.info(t)
            assert info.resident == (1 if path == "resident" else 0)
            got = s.result(t)
            want = oracle_records(r)
            for k in want:
                np.testing.assert_array_equal(got[k], want[k], err_msg=f"{r} {k}")


def test_solo_vs_mix_bit_identical(sv):
    """A mixed trace (both paths, both models, uneven sizes) gives each request
    exactly its solo records, also with a different tick length."""
    trace = sv.make_trace(24, seed=99, n_min=500, n_max=40000, steps_min=20, steps_max=60)
    solo, mix = {}, {}
    with sv.SimulationServer(max_active=1, prefetch=1, sir_resident_max=4096,
                             wealth_resident_max=2048) as s:
        ts = [s.submit(r) for r in trace]
        s.drain(300)
        solo = {i: s.result(t) for i, t in enumerate(ts)}
    with sv.SimulationServer(max_active=16, steps_per_tick=5, sir_resident_max=4096,
                             wealth_resident_max=2048, sir_chunk=4096, wealth_chunk=1024) as s:
        ts = [s.submit(r) for r in trace]
        s.drain(300)
        assert s.stats()["peak_active"] > 1
        mix = {i: s.result(t) for i, t in enumerate(ts)}
    for i in solo:
        assert same(solo[i], mix[i]), trace[i]
    # spot-check two against the oracle
    for i in (0, 5):
        want = oracle_records(trace[i])
        assert same({k: solo[i][k] for k in want}, want)


def test_admission_by_budget(sv):
    """A small arena admits requests as memory frees up; a request larger than
    the budget fails with a message; nothing else is affected."""
    budget = 24 << 20
    trace = sv.make_trace(16, seed=5, n_min=20000, n_max=120000, steps_min=10, steps_max=30)
    with sv.SimulationServer(max_active=16, hbm_budget=budget) as s:
        big = s.submit(sv.SimRequest("sir", n=2_000_000, steps=10, seed=1))
        ts = [s.submit(r) for r in trace]
        s.drain(300)
        assert s.poll(big) == "failed"
        for t in ts:
            assert s.poll(t) == "done"
        st = s.stats()
        assert 0 < st["peak_device_bytes"] <= budget
        for i in (0, 3):
            assert same(s.result(ts[i]), oracle_records(trace[i]))


def test_invalid_requests_fail_cleanly(sv):
    with sv.SimulationServer(max_active=2) as s:
        bad = s.submit(sv.SimRequest("wealth", n=1, steps=5))
        good = s.submit(sv.SimRequest("wealth", n=100, steps=5, seed=3))
        s.drain(60)
        assert s.poll(bad) == "failed"
        assert s.poll(good) == "done"
        with pytest.raises(Exception, match="failed"):
            s.result(bad)


def test_benchmark_trace_prefix_matches_oracle(sv):
    """The first 32 requests of the benchmark's seed-0 trace at full size
    (agent counts up to 247,720, 100-365 steps, both models): through the
    batched server with its tuned configuration and through the plain
    per-request engines, every record equals the oracle's run of the same
    request bit for bit.  (Partitionable layout: request i of a trace does
    not depend on the trace's length, so these are the measured trace's
    first 32.)"""
    import concurrent.futures as cf
    import os

    trace = sv.make_trace(32, seed=0)
    assert sv.make_trace(400, seed=0)[:32] == trace
    assert max(r.n for r in trace) > 200_000 and {r.kind for r in trace} == {"sir", "wealth"}
    with sv.SimulationServer(sv.tuned_config("batched")) as s:
        tickets = [s.submit(r) for r in trace]
        s.drain(300)
        served = [s.result(t) for t in tickets]
    plain = sv.plain_engine_run(trace)["records"]
    threads = max(1, min(16, os.cpu_count() or 1) // 4)

    def want(req):
        if req.kind == "wealth":
            x, y, w, rng = O.wealth_init(req.n, req.seed, O.MODE_PARTITIONABLE)
            mean, gini = O.wealth_run_threads(x, y, w, rng, req.steps, threads,
                                              O.MODE_PARTITIONABLE)
            return {"mean": mean, "gini": gini}
        return oracle_records(req)

    with cf.ThreadPoolExecutor(4) as ex:
        wants = list(ex.map(want, trace))
    for r, a, b, w in zip(trace, served, plain, wants):
        assert same(a, w), r
        assert same(b, w), r
