"""Continuous-batching simulation serving on the HIP server.

Replaces the reference's serial runners -- ``SensitivityAnalysis``'s loop
(jaxabm/analysis.py:97-165), ``run_parallel_simulations`` (jaxabm/utils.py:
128-176), ``Model.batch_run`` (jaxabm/agentpy.py:1116-1152) and the robust
evaluation batches (jaxabm/analysis.py:434-476) -- which run one model after
another.  Here requests of uneven sizes are admitted as device memory and
instance slots free up, kept on the device, and advanced together by
multi-instance step kernels (the native scheduler in ``csrc/serve.hip``).

* :class:`SimulationServer` -- submit / poll / wait / result / drain.
* :func:`make_trace` -- the seeded request trace (Poisson arrivals; wealth or
  SIR; log-uniform agent counts; uniform parameters and step counts), drawn
  from a threefry stream so it is the same on every machine.
* :func:`replay` -- submit a trace at its arrival times (or all at once) and
  measure sims/s and arrival-to-result latency percentiles.
* :func:`autotune` -- the server's own tuner: coordinate search over tick
  length, resident-instance thresholds and range sizes on a sample of the
  trace, stored per configuration in ``serving_tuned.json``.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import threading
import time
from dataclasses import asdict, dataclass, field

import numpy as np

from . import random as jr
from ._lib import (AbmError, ServeConfig, ServeRequest, ServeStats, TicketInfo, check, lib,
                   require_gpu)

WEALTH, SIR = 0, 1
KINDS = {"wealth": WEALTH, "sir": SIR}
STATUS = {0: "queued", 1: "prepared", 2: "uploading", 3: "active", 4: "done", 5: "failed"}
DONE, FAILED = 4, 5
TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "serving_tuned.json")

# algorithmic bytes per agent-step: wealth x,y r/w 16 + wealth r/w 8 +
# scatter r/m/w 8; SIR state r/w 2 + row_ptr 4 + per CSR entry col 4 + state 1
WEALTH_BYTES_PER_AGENT_STEP = 32


@dataclass
class SimRequest:
    kind: str = "wealth"
    n: int = 1000
    steps: int = 100
    seed: int = 0
    graph_seed: int = 0
    m: int = 4
    n_infected: int = 100
    beta: float = 0.05
    gamma: float = 0.1
    arrival: float = 0.0  # seconds after the trace start

    def to_c(self) -> ServeRequest:
        return ServeRequest(KINDS[self.kind], int(self.steps), int(self.n), int(self.seed),
                            int(self.graph_seed), int(self.m), int(self.n_infected),
                            float(self.beta), float(self.gamma))

    def algorithmic_bytes(self) -> int:
        if self.kind == "wealth":
            return WEALTH_BYTES_PER_AGENT_STEP * self.n * self.steps
        nnz = 2 * (self.m * (self.n - self.m) if self.n > self.m else self.n - 1)
        return (6 * self.n + 5 * nnz) * self.steps


def default_config() -> dict:
    c = ServeConfig()
    lib().abm_serve_default_config(C.byref(c))
    return {name: getattr(c, name) for name, _ in ServeConfig._fields_}


def tuned_config(profile: str = "batched") -> dict:
    """Defaults overridden by the autotuner's stored choice for ``profile``."""
    cfg = default_config()
    if profile == "solo":
        cfg.update(max_active=1, prefetch=1)
    try:
        with open(TUNED_PATH) as f:
            cfg.update(json.load(f).get(profile, {}).get("config", {}))
    except (OSError, ValueError):
        pass
    return cfg


class SimulationServer:
    """A running server: one scheduler thread and host preparation workers
    around the device arena."""

    def __init__(self, config: dict | None = None, **overrides):
        require_gpu()
        cfg = default_config()
        cfg.update(config or {})
        cfg.update(overrides)
        self.config = cfg
        c = ServeConfig(**{k: v for k, v in cfg.items()})
        h = C.c_void_p()
        check(lib().abm_serve_create(C.byref(c), C.byref(h)), "serve_create")
        self._h = h
        self._requests: dict[int, SimRequest] = {}

    def close(self) -> None:
        if self._h:
            lib().abm_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, req: SimRequest) -> int:
        t = C.c_int64()
        check(lib().abm_serve_submit(self._h, C.byref(req.to_c()), C.byref(t)), "serve_submit")
        self._requests[t.value] = req
        return t.value

    def poll(self, ticket: int) -> str:
        s = C.c_int32()
        check(lib().abm_serve_poll(self._h, ticket, C.byref(s)), "serve_poll")
        return STATUS[s.value]

    def wait(self, ticket: int, timeout: float = 60.0) -> str:
        s = C.c_int32()
        check(lib().abm_serve_wait(self._h, ticket, float(timeout), C.byref(s)), "serve_wait")
        return STATUS[s.value]

    def drain(self, timeout: float = 600.0) -> None:
        check(lib().abm_serve_drain(self._h, float(timeout)), "serve_drain")

    def info(self, ticket: int) -> TicketInfo:
        info = TicketInfo()
        rc = lib().abm_serve_collect(self._h, ticket, None, 0, C.byref(info))
        if rc and info.status != FAILED:
            check(rc, "serve_collect")
        return info

    def result(self, ticket: int) -> dict:
        """Records of a finished request: wealth {'mean', 'gini'} float32
        [steps]; SIR {'S', 'I', 'R'} int32 [steps]."""
        req = self._requests[ticket]
        info = TicketInfo()
        nbytes = (12 if req.kind == "sir" else 8) * req.steps
        buf = np.empty(nbytes, np.uint8)
        check(lib().abm_serve_collect(self._h, ticket, buf.ctypes.data, nbytes, C.byref(info)),
              "serve_collect")
        if req.kind == "sir":
            c = buf.view(np.int32).reshape(req.steps, 3)
            return {"S": c[:, 0].copy(), "I": c[:, 1].copy(), "R": c[:, 2].copy()}
        f = buf.view(np.float32)
        return {"mean": f[: req.steps].copy(), "gini": f[req.steps:].copy()}

    def stats(self) -> dict:
        st = ServeStats()
        check(lib().abm_serve_get_stats(self._h, C.byref(st)), "serve_stats")
        return {name: getattr(st, name) for name, _ in ServeStats._fields_}


def clock() -> float:
    """The server's clock (steady, seconds): ticket timestamps use it."""
    return float(lib().abm_serve_clock())


# ---------------------------------------------------------------- request trace
def make_trace(n_requests: int, rate: float = 200.0, seed: int = 0, n_min: int = 1000,
               n_max: int = 256000, steps_min: int = 100, steps_max: int = 365,
               wealth_fraction: float = 0.5) -> list[SimRequest]:
    """Seeded request trace: request i takes 8 words of the threefry stream
    split(PRNGKey(seed), 4 n_requests) (partitionable layout) for its
    inter-arrival time (exponential, ``rate`` per second), model kind, agent
    count (log-uniform in [n_min, n_max]), SIR beta ~ U[0.01, 0.1] and gamma ~
    U[0.05, 0.2], and step count ~ U{steps_min..steps_max}.  SIR graphs are
    Barabasi-Albert with m = 4 and 100 initial infections."""
    words = jr.host_split(jr.host_prng_key(seed), 4 * n_requests,
                          jr.MODE_PARTITIONABLE).reshape(n_requests, 8)
    u = (words >> np.uint32(8)).astype(np.float64) / float(1 << 24)  # [0, 1)
    t = 0.0
    out = []
    lo, hi = math.log(n_min), math.log(n_max)
    for i in range(n_requests):
        t += -math.log1p(-u[i, 0]) / rate
        kind = "wealth" if u[i, 1] < wealth_fraction else "sir"
        n = int(math.floor(math.exp(lo + u[i, 2] * (hi - lo))))
        n = max(n_min, min(n, n_max))
        steps = steps_min + int(u[i, 5] * (steps_max - steps_min + 1))
        out.append(SimRequest(kind=kind, n=n, steps=steps, seed=i, graph_seed=1_000_000 + i,
                              m=4, n_infected=100, beta=float(np.float32(0.01 + 0.09 * u[i, 3])),
                              gamma=float(np.float32(0.05 + 0.15 * u[i, 4])), arrival=t))
    return out


def trace_checksum(trace: list[SimRequest]) -> str:
    import hashlib

    h = hashlib.sha256()
    for r in trace:
        h.update(json.dumps(asdict(r), sort_keys=True).encode())
    return h.hexdigest()[:16]


# ---------------------------------------------------------------- measurement
@dataclass
class ReplayReport:
    requests: int
    makespan_s: float
    sims_per_s: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    agent_steps: int
    algorithmic_bytes: int
    hbm_gbps: float
    peak_active: int
    peak_device_bytes: int
    prep_s: float = 0.0       # last request prepared - trace start
    queue_ms: float = 0.0     # mean prepared -> started on the device
    tickets: list = field(default_factory=list, repr=False)
    latencies_ms: list = field(default_factory=list, repr=False)

    def summary(self) -> dict:
        d = asdict(self)
        d.pop("tickets")
        d.pop("latencies_ms")
        return d


def replay(server: SimulationServer, trace: list[SimRequest], time_scale: float = 1.0,
           timeout: float = 600.0) -> ReplayReport:
    """Submit ``trace`` at its arrival times scaled by ``time_scale`` (0: all
    at once) and wait for every result.  Latency = result time - arrival
    time; sims/s = requests / (last result - trace start)."""
    t0 = clock()
    arrivals = [t0 + r.arrival * time_scale for r in trace]
    tickets = []
    for r, ta in zip(trace, arrivals):
        while True:
            dt = ta - clock()
            if dt <= 0:
                break
            time.sleep(min(dt, 0.002) if dt > 0.0005 else 0)
        tickets.append(server.submit(r))
    server.drain(timeout)
    lat, t_end, t_ready, queue = [], t0, t0, []
    for tk, ta in zip(tickets, arrivals):
        info = server.info(tk)
        if info.status != DONE:
            raise AbmError(f"request {tk} did not complete ({STATUS[info.status]})")
        lat.append((info.t_done - ta) * 1e3)
        t_end = max(t_end, info.t_done)
        t_ready = max(t_ready, info.t_ready)
        queue.append((info.t_start - info.t_ready) * 1e3)
    span = max(t_end - t0, 1e-9)
    st = server.stats()
    agent_steps = sum(r.n * r.steps for r in trace)
    alg = sum(r.algorithmic_bytes() for r in trace)
    lat_a = np.array(lat)
    return ReplayReport(
        requests=len(trace), makespan_s=span, sims_per_s=len(trace) / span,
        p50_ms=float(np.percentile(lat_a, 50)), p99_ms=float(np.percentile(lat_a, 99)),
        mean_ms=float(lat_a.mean()), agent_steps=agent_steps, algorithmic_bytes=alg,
        hbm_gbps=alg / span / 1e9, peak_active=st["peak_active"],
        peak_device_bytes=st["peak_device_bytes"], prep_s=t_ready - t0,
        queue_ms=float(np.mean(queue)), tickets=tickets, latencies_ms=lat)


# ---------------------------------------------------------------- autotuner
SEARCH_SPACE = {
    "steps_per_tick": [2, 4, 8, 16, 32, 64],
    "sir_resident_max": [0, 1024, 2048, 4096, 8192, 32768],
    "wealth_resident_max": [0, 512, 1024, 2048, 4096, 16384],
    "sir_chunk": [1024, 2048, 4096, 8192, 16384],
    "wealth_chunk": [128, 256, 512, 1024, 2048, 4096],
}


def _measure(cfg: dict, trace: list[SimRequest]) -> float:
    with SimulationServer(cfg) as s:
        return replay(s, trace, time_scale=0.0).sims_per_s


def autotune(profile: str = "batched", n_requests: int = 192, seed: int = 12345,
             rounds: int = 2, write: bool = True, log=print) -> dict:
    """Coordinate search over ``SEARCH_SPACE`` maximising saturated sims/s on a
    sample trace (its own seed, not the benchmark's).  ``profile`` "solo"
    tunes the one-at-a-time server (max_active = 1) the same way, so the
    baseline runs its best configuration too."""
    trace = make_trace(n_requests, seed=seed)
    cfg = tuned_config(profile)
    best = _measure(cfg, trace)
    log(f"[autotune {profile}] start {best:.1f} sims/s")
    for _ in range(rounds):
        changed = False
        for key, values in SEARCH_SPACE.items():
            for v in values:
                if v == cfg[key]:
                    continue
                trial = dict(cfg, **{key: v})
                r = _measure(trial, trace)
                log(f"[autotune {profile}] {key}={v}: {r:.1f} sims/s")
                if r > best * 1.02:
                    best, cfg, changed = r, trial, True
        if not changed:
            break
    chosen = {k: cfg[k] for k in SEARCH_SPACE}
    if write:
        data = {}
        try:
            with open(TUNED_PATH) as f:
                data = json.load(f)
        except (OSError, ValueError):
            pass
        data[profile] = {"config": chosen, "sims_per_s": best, "sample_requests": n_requests,
                         "sample_seed": seed}
        with open(TUNED_PATH, "w") as f:
            json.dump(data, f, indent=2, sort_keys=True)
    return chosen


def _thread_submit(server, trace, out):
    out.extend(server.submit(r) for r in trace)


def submit_concurrently(server: SimulationServer, traces: list[list[SimRequest]]) -> list:
    """Submit several traces from several threads at once (the C-ABI submit is
    thread-safe); returns the tickets per trace."""
    outs = [[] for _ in traces]
    threads = [threading.Thread(target=_thread_submit, args=(server, t, o))
               for t, o in zip(traces, outs)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    return outs


def _records_equal(a: dict, b: dict) -> bool:
    return a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a)


def plain_engine_run(trace: list[SimRequest], mode: int | None = None) -> dict:
    """The plain per-run GPU baseline: each request alone, one after another,
    through the single-instance engines (``SIRNetworkEngine``, whose step
    kernel skips quiet rows and agents; ``WealthEngine``, one workgroup per
    run with every draw made before the run).  Per request the timing covers
    the engine's device set-up (graph upload, chunking, seeding), the run and
    the copy of its records to the host; each SIR graph's host CSR is built
    before the clock starts, so this baseline leaves out work the server's
    timing includes.  Latencies at the trace's arrival times follow from the
    measured service times of one in-order executor (a single-server queue)."""
    import torch

    from .network import ba_csr
    from .programs.sir import SIRNetworkEngine
    from .programs.wealth import WealthEngine

    graphs = {i: ba_csr(r.n, r.m, r.graph_seed) for i, r in enumerate(trace) if r.kind == "sir"}
    records, service = [], []
    torch.cuda.synchronize()
    for i, r in enumerate(trace):
        t0 = time.perf_counter()
        if r.kind == "sir":
            eng = SIRNetworkEngine(r.n, csr=graphs[i], beta=r.beta, gamma=r.gamma,
                                   n_seed_infected=r.n_infected, seed=r.seed, mode=mode)
            c = eng.run(r.steps)
            rec = {"S": c[:, 0].copy(), "I": c[:, 1].copy(), "R": c[:, 2].copy()}
        else:
            eng = WealthEngine(r.n, seed=r.seed, mode=mode)
            mean, gini = eng.run(r.steps)
            rec = {"mean": mean, "gini": gini}
        service.append(time.perf_counter() - t0)
        records.append(rec)
        del eng
    total = float(np.sum(service))
    end, lat = 0.0, []
    for r, dt in zip(trace, service):
        start = max(end, r.arrival)
        end = start + dt
        lat.append((end - r.arrival) * 1e3)
    lat_a = np.array(lat)
    return {"records": records, "sims_per_s": len(trace) / total, "total_s": total,
            "p50_ms": float(np.percentile(lat_a, 50)), "p99_ms": float(np.percentile(lat_a, 99)),
            "mean_service_ms": total / len(trace) * 1e3}


def benchmark(n_requests: int = 400, rate: float = 200.0, seed: int = 0,
              solo: dict | None = None, batched: dict | None = None, log=None,
              plain: bool = True) -> dict:
    """The serving workload on one trace: one request at a time (max_active =
    1, its own tuned configuration) against continuous batching.

    For each server: a saturated replay (every request submitted at once:
    sims/s is the capacity) and a replay at the trace's arrival times
    (``rate`` per second: latency percentiles arrival -> result).  Every
    request's records are compared between the two servers."""
    trace = make_trace(n_requests, rate=rate, seed=seed)
    out = {"requests": n_requests, "rate": rate, "trace_seed": seed,
           "trace_checksum": trace_checksum(trace)}
    results = {}
    for name, cfg in (("solo", solo or tuned_config("solo")),
                      ("batched", batched or tuned_config("batched"))):
        with SimulationServer(cfg) as s:
            sat = replay(s, trace, time_scale=0.0)
            results[name] = [s.result(t) for t in sat.tickets]
        with SimulationServer(cfg) as s:
            live = replay(s, trace, time_scale=1.0)
        out[name] = {"config": {k: cfg[k] for k in SEARCH_SPACE if k in cfg} |
                     {"max_active": cfg["max_active"]},
                     "saturated": sat.summary(), "at_rate": live.summary()}
        if log:
            log(f"[serving] {name}: saturated {sat.sims_per_s:.1f} sims/s, at {rate:g}/s "
                f"p50 {live.p50_ms:.2f} ms p99 {live.p99_ms:.2f} ms")
    out["identical"] = all(_records_equal(a, b) for a, b in zip(results["solo"],
                                                                 results["batched"]))
    out["speedup"] = (out["batched"]["saturated"]["sims_per_s"] /
                      out["solo"]["saturated"]["sims_per_s"])
    if plain:
        cfg = batched or tuned_config("batched")
        p = plain_engine_run(trace, mode=int(cfg.get("mode", jr.MODE_PARTITIONABLE)))
        out["plain_engine"] = {k: v for k, v in p.items() if k != "records"}
        out["identical_plain"] = all(_records_equal({k: a[k] for k in b}, b)
                                     for a, b in zip(results["batched"], p["records"]))
        out["speedup_vs_plain"] = out["batched"]["saturated"]["sims_per_s"] / p["sims_per_s"]
        if log:
            log(f"[serving] plain engines: {p['sims_per_s']:.1f} sims/s, p50 {p['p50_ms']:.2f} "
                f"ms p99 {p['p99_ms']:.2f} ms")
    return out
