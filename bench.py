#!/usr/bin/env python3
"""Headline benchmark: agent-steps/s of the 1M-agent network SIR program.

Workload (one bench "step" = one complete simulation from the seeded initial
state): Barabasi-Albert graph n = 1,000,000, m = 4 (E = 3,999,984 undirected
edges, int32 CSR in both directions), uint8 S/I/R state, 100 seeded
infections, beta = 0.05, gamma = 0.1, 365 synchronous steps, S/I/R counts
recorded every step, JAX partitionable threefry stream, seed 0.  The run is a
captured hipGraph (state reset copy + counts clear + 365 step kernels) with
all inputs resident in HBM.

Also reported:
  * roofline of the step kernel: algorithmic bytes per launch (46 B/agent-step
    x 1M = 45,999,844 B: state read + write, row_ptr, col, one neighbour-state
    byte per CSR entry) over the kernel's average duration, taken as the
    HIP-event time of the timed graph runs on their stream divided by the
    number of step kernels, against MI355X's 8 TB/s;
  * an upper bound of the data-dependent bytes the skipping kernel touches;
  * parity: the graph's CSR checksum and the recorded S/I/R rows must equal
    the committed fixture (tests/golden/sir_ba1M_m4_365steps.json: the oracle's
    run on the same graph), or nothing is reported and the exit status is 3;
  * cpu_baseline: the C oracle (a restatement of the reference semantics, one
    thread) running the same full simulation on the host;
  * serving: the batched-serving workload (seeded trace of wealth-exchange and
    SIR requests, agent counts log-uniform in [1k, 256k], 100-365 steps,
    Poisson arrivals at 200/s) through the continuous-batching server against
    the same server running one request at a time: saturated sims/s of each,
    arrival-to-result p50/p99 at 200 requests/s, and whether every request's
    records are identical between the two;
  * wealth, schelling, deffuant: the other programs' lines (1,000-agent
    100-step wealth exchange, 512 x 512 200-step Schelling, 65,536-agent x
    256-instance 200-step Deffuant gradient evaluation), each with its
    cpu_baseline and roofline;
  * generic: a 10,000-agent 200-step model with no library kernel, run
    eagerly and as graph replays, vmapped and fused (steps/s of each, the
    value being graph replays of the fused update).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import subprocess
import sys
import time

N_AGENTS = 1_000_000
BA_M = 4
SIM_STEPS = 365
BETA = 0.05
GAMMA = 0.1
N_SEED = 100
SEED = 0
GRAPH_SEED = 0
HBM_PEAK_GBS = 8000.0
ALGO_BYTES_PER_AGENT_STEP = 46.0
MARK_HORIZON = 16  # the step kernel's mark horizon (steps a mark stays valid)


def cpu_model() -> str:
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True).stdout
        for line in out.splitlines():
            if line.startswith("Model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def algorithmic_bytes(n: int, nnz: int) -> int:
    # state read n + state write n + row_ptr 4(n+1) + col 4 nnz + gather 1 nnz
    return 2 * n + 4 * (n + 1) + 5 * nnz


def cpu_baseline(row_ptr, col, runs: int) -> dict:
    """The C oracle (a restatement of the reference semantics, not JAX) on
    the full 365-step 1M-agent simulation: median of ``runs`` runs after one
    warm-up, single-threaded (the reported value) and on all host cores."""
    import numpy as np

    import oracle as O

    def timed(threads):
        times = []
        for i in range(runs + 1):
            st, rng = O.sir_init(N_AGENTS, N_SEED, SEED, O.MODE_PARTITIONABLE)
            t0 = time.perf_counter()
            O.sir_run_threads(row_ptr, col, st, rng, BETA, GAMMA, SIM_STEPS, threads,
                              O.MODE_PARTITIONABLE)
            if i:
                times.append(time.perf_counter() - t0)
        return float(np.median(times))

    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    one = timed(1)
    many = timed(ncores)
    return {
        "value": N_AGENTS * SIM_STEPS / one,
        "unit": "agent-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"full 365-step 1M-agent SIR simulation (the measured workload itself), C "
                   f"oracle: CPU restatement of the reference semantics, not JAX; median of "
                   f"{runs} runs after a warm-up; host CPU: {cpu_model()}, nproc={os.cpu_count()}"),
        "run_s": one,
        "all_cores": {"value": N_AGENTS * SIM_STEPS / many, "cores": ncores, "run_s": many,
                      "kind": "port (OpenMP over the agents of each step)"},
    }


def cpu_sample(fn, units: float, unit: str, sample: str, runs: int) -> dict:
    """``fn`` (one bounded run of the C oracle, a restatement of the
    reference semantics, single-threaded) timed as the median of ``runs``
    runs after a warm-up; ``units`` of work per run."""
    import numpy as np

    times = []
    for i in range(runs + 1):
        t0 = time.perf_counter()
        fn()
        if i:
            times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": units / t, "unit": unit, "cores": 1, "kind": "port",
            "sample": f"{sample}; C oracle, not JAX; median of {runs} runs after a warm-up; "
                      f"host CPU: {cpu_model()}, nproc={os.cpu_count()}",
            "run_s": t}


PROFILES = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles")


def measured_traffic(kernel: str = "k_sir_step"):
    """HBM bytes per launch of ``kernel`` from the committed rocprofv3 PMC
    summary profiles/<kernel>_traffic.json (FETCH_SIZE + WRITE_SIZE per
    dispatch, scaled by the same counters' reading for a copy of known size),
    or None."""
    path = os.path.join(PROFILES, f"{kernel}_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return {"bytes_per_launch": d["bytes_per_launch"], "source": "profiles/" +
                os.path.basename(path)}
    except (OSError, ValueError, KeyError):
        return None


def pmc_probe(workload: str = "sir") -> int:
    """A 256 MiB device-to-device copy (3 times: the counters' calibration),
    then one 365-step SIR simulation launched step by step, or (deffuant) one
    forward + adjoint gradient evaluation of the 65,536 x 256 x 200 batch."""
    import torch

    nbytes = 256 << 20
    src = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    if workload == "deffuant":
        import numpy as np

        from jaxabm_amd.programs.deffuant import DeffuantModel, target_stats

        n, B, S, K = 65536, 256, 200, 20
        target = target_stats(n, S, seed=0)
        g = np.linspace(0.05, 0.5, 16, dtype=np.float32)
        m = DeffuantModel(n, B, S, seed=0, ckpt_every=K)
        loss = m.forward(np.repeat(g, 16), np.tile(g, 16), target)
        ge, gm = m.backward()
        torch.cuda.synchronize()
        print(json.dumps({"probe": "copy 256 MiB x3 (read 256 MiB + write 256 MiB each) + one "
                          "Deffuant forward + adjoint (65,536 x 256 x 200, checkpoints every 20)",
                          "copy_bytes": nbytes, "loss_sum": float(loss.sum()),
                          "grad_abs_sum": float(ge.abs().sum() + gm.abs().sum())}))
        return 0
    from jaxabm_amd.network import ba_csr
    from jaxabm_amd.programs.sir import SIRNetworkEngine

    row_ptr, col = ba_csr(N_AGENTS, BA_M, GRAPH_SEED)
    eng = SIRNetworkEngine(N_AGENTS, csr=(row_ptr, col), beta=BETA, gamma=GAMMA,
                           n_seed_infected=N_SEED, seed=SEED)
    counts = eng.run(SIM_STEPS)
    print(json.dumps({"probe": "copy 256 MiB x3 (read 256 MiB + write 256 MiB each) + one "
                      "365-step SIR simulation", "copy_bytes": nbytes,
                      "final_counts": counts[-1].tolist()}))
    return 0


def _timed(stream, fn, runs: int) -> float:
    """Seconds per call of fn(), HIP events on ``stream`` around ``runs`` calls."""
    import torch

    from jaxabm_amd._lib import check, lib

    sh = C.c_void_p(stream.cuda_stream)
    torch.cuda.synchronize()
    check(lib().abm_event_timer(sh, 0, None), "timer")
    for _ in range(runs):
        fn()
    ms = C.c_float()
    check(lib().abm_event_timer(sh, 1, C.byref(ms)), "timer")
    torch.cuda.synchronize()
    return ms.value / 1e3 / runs


def copy_reference_gbs(stream) -> dict:
    """Device-to-device copy rate of 1 GiB (read + write bytes / time), the
    plain streaming ceiling the kernels' HBM rates compare with; a copy this
    size does not fit the 256 MB Infinity Cache."""
    import torch

    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    t = _timed(stream, lambda: dst.copy_(src), 20)
    del src, dst
    return {"bytes": 2 << 30, "GBs": (2 << 30) / t / 1e9, "frac": (2 << 30) / t / 1e9 / HBM_PEAK_GBS,
            "what": "torch copy_ of 1 GiB, read + write bytes per second"}


def bench_deffuant(args) -> dict:
    """Differentiable Deffuant: 65,536 agents x 256 (eps, mu) instances x 200
    steps, checkpoints every 20; one bench step = one forward + adjoint
    gradient evaluation of the whole batch."""
    import numpy as np
    import torch

    from jaxabm_amd.programs.deffuant import DeffuantModel, target_stats

    n, B, S, K = 65536, 256, 200, 20
    target = target_stats(n, S, seed=0)
    g = np.linspace(0.05, 0.5, 16, dtype=np.float32)
    eps, mu = np.repeat(g, 16), np.tile(g, 16)
    m = DeffuantModel(n, B, S, seed=0, ckpt_every=K)
    stream = torch.cuda.current_stream()

    def fwd():
        m.forward(eps, mu, target)

    def both():
        m.forward(eps, mu, target)
        m.backward()

    for _ in range(args.warmup):
        both()
    t_fwd = _timed(stream, fwd, args.steps)
    t_both = _timed(stream, both, args.steps)
    m.forward(eps, mu, target)
    g20 = m.backward()
    full_tape = _deffuant_full_tape(n, B, S, eps, mu, target, g20, stream, args)
    del m
    torch.cuda.empty_cache()
    fwd_bytes = n * B * 8 + 4 * n  # x read + write, the pairing
    step_s = t_fwd / S
    achieved = fwd_bytes / step_s / 1e9
    traffic = measured_traffic("k_deffuant_step")
    return {
        "metric": "agent-instance-steps/s (Deffuant gradient evaluation, forward + adjoint)",
        "value": n * B * S / t_both, "unit": "agent-instance-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t_both * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded opinions, 16 x 16 grid of (eps, mu) in [0.05, 0.5]^2)",
        "config": {"workload": "deffuant_65536x256_200steps_ckpt20", "n_agents": n,
                   "instances": B, "sim_steps": S, "ckpt_every": K,
                   "bench_step": "one forward + adjoint of the batch"},
        "forward_ms": t_fwd * 1e3, "adjoint_ms": (t_both - t_fwd) * 1e3,
        "full_tape": full_tape,
        "cpu_baseline": _deffuant_cpu(n, eps, mu, target, args.cpu_runs),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic["bytes_per_launch"] if traffic else None,
                     "traffic_source": traffic["source"] if traffic else None,
                     "kernel": "k_deffuant_step (forward average incl. final statistics)",
                     "algorithmic_bytes_per_launch": fwd_bytes, "avg_kernel_us": step_s * 1e6,
                     "copy_reference": copy_reference_gbs(stream)},
    }


def _deffuant_full_tape(n, B, S, eps, mu, target, g20, stream, args) -> dict:
    """The same gradient evaluation keeping every step's state in HBM
    (ckpt_every = 1: 13.4 GB for 65,536 x 256 x 200, which 288 GB holds), so
    the adjoint recomputes nothing; its gradients must equal the
    checkpointed run's bit for bit.  Reported beside the measured
    configuration (checkpoints every 20 steps), not as its value."""
    import torch

    from jaxabm_amd.programs.deffuant import DeffuantModel

    m = DeffuantModel(n, B, S, seed=0, ckpt_every=1)

    def both():
        m.forward(eps, mu, target)
        return m.backward()

    for _ in range(max(args.warmup, 1)):
        g1 = both()
    t = _timed(stream, both, args.steps)
    same = all(torch.equal(a, b) for a, b in zip(g1, g20))
    tape_gb = m.ckpt.numel() * 4 / 1e9
    del m
    torch.cuda.empty_cache()
    return {"ckpt_every": 1, "tape_GB": tape_gb, "ms_per_step": t * 1e3,
            "value": n * B * S / t, "unit": "agent-instance-steps/s",
            "grads_identical_to_ckpt20": same,
            "note": "every step's opinions kept in HBM: the adjoint reads them instead of "
                    "recomputing 9 of the 10 checkpoint segments"}


def _deffuant_cpu(n, eps, mu, target, runs):
    if runs <= 0:
        return None
    import numpy as np

    import oracle as O

    b, steps, mode = 32, 10, O.MODE_PARTITIONABLE
    x0, rng = O.deffuant_init(n, b, 0, mode)
    e, m = eps[:b].astype(np.float64), mu[:b].astype(np.float64)
    out = cpu_sample(lambda: O.deffuant_grad(x0, e, m, rng, steps, target, mode), n * b * steps,
                     "agent-instance-steps/s",
                     f"forward + float64 reverse sweep of {n:,} agents x {b} of the 256 "
                     f"(eps, mu) instances x {steps} steps (the per-step pairing is shared by "
                     f"the instances, as in the measured batch)", min(runs, 3))
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    many = cpu_sample(lambda: O.deffuant_grad_threads(x0, e, m, rng, steps, target, ncores, mode),
                      n * b * steps, "agent-instance-steps/s", "", min(runs, 3))
    out["all_cores"] = {"value": many["value"], "cores": ncores, "run_s": many["run_s"],
                        "kind": "port (OpenMP over the instances; same bits)"}
    return out


def _wealth_cpu(n, steps, runs):
    if runs <= 0:
        return None
    import oracle as O

    mode = O.MODE_PARTITIONABLE

    def run():
        x, y, w, rng = O.wealth_init(n, 0, mode)
        O.wealth_run(x, y, w, rng, steps, mode)

    return cpu_sample(run, steps, "steps/s",
                      f"the full {steps}-step {n:,}-agent simulation (init included, as on the "
                      f"GPU)", max(runs, 5))


def _schelling_cpu(h, w, runs):
    if runs <= 0:
        return None
    import oracle as O

    mode, steps = O.MODE_PARTITIONABLE, 20
    init = O.schelling_init(h, w, 0, mode)

    def run():
        cells, pos, typ, rng = (a.copy() for a in init)
        O.schelling_run(h, w, cells, pos, typ, rng, steps, mode)

    out = cpu_sample(run, h * w * steps, "cell-steps/s",
                     f"the first {steps} of the 200 steps (the busiest) of the {h}x{w} "
                     f"simulation from the same seeded grid", min(runs, 3))
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()

    def run_all():
        cells, pos, typ, rng = (a.copy() for a in init)
        O.schelling_run_threads(h, w, cells, pos, typ, rng, steps, ncores, mode)

    many = cpu_sample(run_all, h * w * steps, "cell-steps/s", "", min(runs, 3))
    out["all_cores"] = {"value": many["value"], "cores": ncores, "run_s": many["run_s"],
                        "kind": "port (OpenMP over the agents of each step; claims and moves "
                                "sequential; same results)"}
    return out


def _trader_cpu(n, steps, runs):
    if runs <= 0:
        return None
    import oracle as O

    mode = O.MODE_PARTITIONABLE

    def run():
        x, w, rng = O.trader_init(n, 0, mode)
        O.trader_run(x, w, rng, steps, mode)

    return cpu_sample(run, steps, "steps/s",
                      f"the full {steps}-step {n:,}-agent model (the oracle's restatement of its "
                      f"agent code and key schedule)", min(runs, 3))


def bench_wealth(args) -> dict:
    """Wealth exchange, 1,000 agents, 100 steps, seed 0: one bench step = one
    whole simulation (state init + one single-workgroup launch)."""
    import numpy as np
    import torch

    from jaxabm_amd._lib import check, lib, ptr, stream_handle
    from jaxabm_amd.programs.wealth import WealthEngine

    n, S = 1000, 100
    eng = WealthEngine(n, seed=0)
    keys = eng.key_table(S)
    mean = torch.empty(S, dtype=torch.float32, device="cuda")
    gini = torch.empty(S, dtype=torch.float32, device="cuda")
    draws = eng.draw_buffer(S)
    desc = torch.from_numpy(eng.descriptor(keys, S, mean, gini, draws).view(np.uint8)).to("cuda")
    from jaxabm_amd import random as jr

    ik = torch.from_numpy(jr.host_split(jr.host_prng_key(0), 2, eng.mode)[1].view(np.int32)
                          .copy()).to("cuda")
    stream = torch.cuda.current_stream()

    def run():
        check(lib().abm_wealth_init(ptr(ik), n, ptr(eng.x), ptr(eng.y), ptr(eng.w), eng.mode,
                                    stream_handle()), "init")
        check(lib().abm_wealth_run_batch(ptr(desc), 1, eng.n, eng.mode, stream_handle()), "run")

    for _ in range(args.warmup):
        run()
    t = _timed(stream, run, args.steps)
    return {
        "metric": "steps/s (wealth exchange, 1,000 agents)", "value": S / t, "unit": "steps/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "i32",
        "data": "synthetic (seeded)",
        "config": {"workload": "wealth_1000agents_100steps", "n_agents": n, "sim_steps": S,
                   "bench_step": "one full simulation"},
        "run_latency_us": t * 1e6, "agent_steps_per_s": n * S / t,
        "cpu_baseline": _wealth_cpu(n, S, args.cpu_runs),
        "roofline": {"bound": "latency", "achieved": None, "peak": None, "unit": None,
                     "frac": None, "traffic": None,
                     "note": "one workgroup holds the 32 KB/step working set; launch- and "
                             "barrier-bound, a bandwidth roofline does not apply"},
    }


class _Trader:
    """A generic agent written against the reference API (per-agent update
    of torch ops under vmap): a position walk with a normal draw and a
    coin-flip wealth change."""

    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        return {"x": jr.uniform(k1), "w": jr.randint(k2, (), 0, 6)}

    def update(self, state, model_state, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        x = jnp.mod(state["x"] + jr.normal(k1) * 0.01, 1.0)
        give = jr.uniform(k2) < 0.5
        return {"x": x, "w": jnp.where(give & (state["w"] > 0), state["w"] - 1, state["w"] + 1)}


def _trader_env(env, agents, params, key):
    from jaxabm_amd import numpy as jnp

    new = dict(env)
    new["mean_x"] = jnp.mean(agents["t"]["x"])
    return new


def _trader_metrics(env, agents, params):
    from jaxabm_amd import numpy as jnp

    a = agents["t"]
    return {"mean_x": env["mean_x"], "total_w": jnp.sum(a["w"]), "rich": jnp.sum(a["w"] > 3)}


def bench_generic(args) -> dict:
    """A generic model (no library kernel): 10,000 agents, 200 steps, per-agent
    torch ops, an env callback and three record-time reductions.  One bench
    step = one 200-step run.  Four ways, same results (compared bit for bit):
    eager loop or graph replays (Model.run(graph=True)), each with the agent
    update vmapped (one kernel per primitive) or traced into one generated
    kernel (jaxabm_amd.fuse, the default); ``value`` is graph + fused."""
    import contextlib
    import io

    import numpy as np
    import torch

    import jaxabm_amd as J
    from jaxabm_amd import fuse

    n, S = 10_000, 200

    def make():
        m = J.JaxModel(params={}, config=J.ModelConfig(steps=S, seed=0),
                       update_state_fn=_trader_env, metrics_fn=_trader_metrics)
        m.add_agent_collection("t", J.AgentCollection(_Trader(), n))
        m.initialize()
        return m

    stream = torch.cuda.current_stream()
    runs = max(1, min(args.steps, 5))
    out, models = {}, {}
    quiet = io.StringIO()
    for fused in (False, True):
        fuse.enable(fused)
        try:
            for graph in (False, True):
                m = make()
                with contextlib.redirect_stdout(quiet):
                    for _ in range(args.warmup):
                        m.run(S, graph=graph)
                    t = _timed(stream, lambda: m.run(S, graph=graph), runs)
                name = ("graph" if graph else "eager") + ("_fused" if fused else "_vmap")
                out[name] = S / t
                models[name] = m
        finally:
            fuse.enable(True)
    ref = models["eager_vmap"]
    same = all(
        all(torch.equal(ref.agent_collections["t"].states[k], m.agent_collections["t"].states[k])
            for k in ("x", "w")) and bool(np.array_equal(ref._rng, m._rng))
        for m in models.values())
    t_best = S / out["graph_fused"]
    return {
        "metric": "steps/s (generic model, 10,000 agents, graph replays of the fused update)",
        "value": out["graph_fused"], "unit": "steps/s", "n_gpus": 1, "steps": runs,
        "warmup": args.warmup, "ms_per_step": t_best * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+i32", "data": "synthetic (seeded)",
        "config": {"workload": "generic_trader_10000agents_200steps", "n_agents": n,
                   "sim_steps": S, "bench_step": "one 200-step Model.run"},
        "steps_per_s": out, "speedup_vs_eager_vmap": out["graph_fused"] / out["eager_vmap"],
        "cpu_baseline": _trader_cpu(n, S, args.cpu_runs),
        "identical_results": same, "graph_info": models["graph_fused"].last_run_info,
        "fuse": fuse.stats(),
    }


def bench_schelling(args) -> dict:
    """Schelling, 512 x 512 torus, 200 steps: one bench step = one whole
    simulation from the seeded grid (the seeded grid, agent ids and positions
    restored by device copies, then the 200 steps), replayed as one hipGraph
    like the SIR line."""
    import numpy as np
    import torch

    from jaxabm_amd._lib import check, lib, ptr
    from jaxabm_amd.programs.schelling import SchellingEngine

    H = W = 512
    S = 200
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = SchellingEngine(H, W, seed=0)
        keys = eng.key_table(S)
        counters = torch.empty((S, 2), dtype=torch.int32, device="cuda")
        happy = torch.empty(S, dtype=torch.float32, device="cuda")
        moves = torch.empty(S, dtype=torch.int32, device="cuda")
        init = {k: getattr(eng, k).clone() for k in ("cells", "id", "pos")}
    stream.synchronize()
    sh = C.c_void_p(stream.cuda_stream)

    def enqueue_run():
        for k, t in init.items():
            dst = getattr(eng, k)
            check(lib().abm_memcpy_d2d(ptr(dst), ptr(t), C.c_size_t(t.numel() * t.element_size()),
                                       sh), "reset")
        eng.launch(keys, S, counters, happy, moves, stream=stream)

    enqueue_run()  # reference records, launched directly
    stream.synchronize()
    ref = (happy.cpu().numpy().copy(), moves.cpu().numpy().copy())
    g = C.c_void_p()
    check(lib().abm_graph_begin(sh), "graph_begin")
    enqueue_run()
    check(lib().abm_graph_end(sh, C.byref(g)), "graph_end")

    def run_one_graph():
        check(lib().abm_graph_launch(g, sh), "graph_launch")

    for _ in range(args.warmup):
        run_one_graph()
    t_one = _timed(stream, run_one_graph, args.steps)
    same = bool(np.array_equal(happy.cpu().numpy(), ref[0]) and
                np.array_equal(moves.cpu().numpy(), ref[1]))
    check(lib().abm_graph_destroy(g), "graph_destroy")

    # the measured run: launches of 16 steps, each replayed as its own graph
    # (the state restore in the first), stopped once the grid has settled
    chunk = SchellingEngine.SETTLE_CHUNK
    graphs = []
    for s0 in range(0, S, chunk):
        s1 = min(s0 + chunk, S)
        gi = C.c_void_p()
        check(lib().abm_graph_begin(sh), "graph_begin")
        if s0 == 0:
            for k, t0 in init.items():
                check(lib().abm_memcpy_d2d(ptr(getattr(eng, k)), ptr(t0),
                                           C.c_size_t(t0.numel() * t0.element_size()), sh),
                      "reset")
        eng.launch(keys[s0:], s1 - s0, counters[s0:], happy[s0:], moves[s0:], stream=stream,
                   prev=counters[s0 - 1] if s0 else None)
        check(lib().abm_graph_end(sh, C.byref(gi)), "graph_end")
        graphs.append(gi)
    launched = []

    def run():
        launched.append(eng.run_settling(
            keys, S, counters, happy, moves, chunk, stream,
            launch=lambda i, s0, s1: check(lib().abm_graph_launch(graphs[i], sh), "graph_launch")))

    with torch.cuda.stream(stream):
        happy.fill_(-1.0)   # the measured runs must write every record again
        moves.fill_(-1)
    for _ in range(args.warmup):
        run()
    t = _timed(stream, run, args.steps)
    same = same and bool(np.array_equal(happy.cpu().numpy(), ref[0]) and
                         np.array_equal(moves.cpu().numpy(), ref[1]))
    for gi in graphs:
        check(lib().abm_graph_destroy(gi), "graph_destroy")
    # identity steps (the previous step found every agent happy) read one
    # count and move nothing: only computed steps are charged the 10.8 B
    hf = np.asarray(ref[0])
    computed = int(sum(1 for i in range(S) if i == 0 or hf[i - 1] < 1.0))
    cell_bytes = 10.8
    achieved = cell_bytes * H * W * computed / t / 1e9
    return {
        "metric": "cell-steps/s (Schelling 512 x 512)", "value": H * W * S / t,
        "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "i8", "data": "synthetic (seeded grid)",
        "config": {"workload": "schelling_512x512_200steps", "cells": H * W, "sim_steps": S,
                   "hipgraph": True,
                   "bench_step": "one full simulation from the seeded grid (state restored "
                                 "by device copies, then 200 steps)",
                   "launch": f"{chunk} steps per graph replay; after each the host reads the "
                             "last step's happy count back, and once a launch ends with every "
                             "agent happy the steps after the one queued behind it are "
                             "recorded as identity steps without a launch"},
        "replays_identical": same,
        "steps_launched": launched[-1],
        "one_graph_ms_per_run": t_one * 1e3,
        "one_graph_note": "all 200 steps as one graph replay (identity steps launched, each "
                          "exiting after reading the previous step's count)",
        "cpu_baseline": _schelling_cpu(H, W, args.cpu_runs),
        "final_happy_fraction": float(ref[0][-1]), "total_moves": int(ref[1].sum()),
        "computed_steps": computed, "identity_steps": S - computed,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "note": "10.8 algorithmic B per cell-step of the computed steps (2 "
                             "launches each) over the whole run's time; identity steps after "
                             "the grid settles move nothing and are not charged; the 2.8 MB/step "
                             "working set is cache-resident, so this is launch/barrier-bound"},
    }


def serving_summary(n_requests: int) -> dict:
    from jaxabm_amd import serving

    r = serving.benchmark(n_requests=n_requests, rate=200.0, seed=0)
    b, s = r["batched"], r["solo"]
    return {
        "metric": "sims/s, batched serving vs one request at a time (the server at max_active 1, "
                  "and plain single-instance engines)",
        "sims_per_s": b["saturated"]["sims_per_s"],
        "solo_sims_per_s": s["saturated"]["sims_per_s"],
        "speedup": r["speedup"],
        "identical_results": r["identical"],
        "p50_ms": b["at_rate"]["p50_ms"],
        "p99_ms": b["at_rate"]["p99_ms"],
        "solo_p50_ms": s["at_rate"]["p50_ms"],
        "solo_p99_ms": s["at_rate"]["p99_ms"],
        "arrival_rate": r["rate"],
        "requests": r["requests"],
        "trace_checksum": r["trace_checksum"],
        "algorithmic_GBs": b["saturated"]["hbm_gbps"],
        "peak_active": b["saturated"]["peak_active"],
        "batched_config": b["config"],
        "solo_config": s["config"],
        "plain_engine_sims_per_s": r["plain_engine"]["sims_per_s"],
        "speedup_vs_plain_engine": r["speedup_vs_plain"],
        "identical_results_plain_engine": r["identical_plain"],
        "plain_engine_p50_ms": r["plain_engine"]["p50_ms"],
        "plain_engine_p99_ms": r["plain_engine"]["p99_ms"],
        "plain_engine_note": ("each request alone through SIRNetworkEngine / WealthEngine, one "
                              "after another; device set-up, run and result copy timed, host "
                              "BA graph build excluded; latency from a single-server queue over "
                              "the measured service times"),
    }


def bench_serving(args) -> dict:
    """The batched-serving line alone: the seeded trace of ``--serving-requests``
    requests, saturated and at 200 requests/s, against the server at
    max_active 1 and the plain engines."""
    srv = serving_summary(args.serving_requests)
    return {
        "metric": "sims/s (batched serving of a seeded SIR + wealth request trace)",
        "value": srv["sims_per_s"], "unit": "sims/s", "n_gpus": 1,
        "steps": srv["requests"], "warmup": 0, "ms_per_step": 1e3 / srv["sims_per_s"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u8+i32",
        "data": "synthetic (seeded request trace)",
        "config": {"workload": f"serving_trace_{srv['requests']}requests",
                   "trace_checksum": srv["trace_checksum"],
                   "bench_step": "one simulation request"},
        "serving": srv,
    }


def sir_fixture() -> dict:
    """The committed checksum of the measured graph and the oracle's S/I/R
    rows of the measured run (tests/golden/sir_ba1M_m4_365steps.json)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden",
                        "sir_ba1M_m4_365steps.json")
    with open(path) as f:
        return json.load(f)


class ParityError(RuntimeError):
    """The measured workload is not the parity-tested one."""


def bench_sir(args) -> dict:
    """The headline: 1M-agent BA SIR, 365 steps, one hipGraph replay per
    bench step.  The graph must have the committed checksum and the recorded
    S/I/R rows must equal the oracle's (both from the committed fixture), or
    ParityError is raised before anything is reported."""
    import numpy as np
    import torch

    from jaxabm_amd._lib import check, lib, ptr
    from jaxabm_amd.network import ba_csr_device, csr_checksum
    from jaxabm_amd.programs.sir import SIRNetworkEngine

    gold = sir_fixture()
    # the graph is built on the device (BA edges, then the CSR), outside the
    # timed region; its host copy is checksummed.  A small build first loads
    # the kernels, so graph_build_ms is the build itself.
    ba_csr_device(1000, BA_M, GRAPH_SEED)
    torch.cuda.synchronize()
    t_build = time.perf_counter()
    net = ba_csr_device(N_AGENTS, BA_M, GRAPH_SEED)
    torch.cuda.synchronize()
    build_ms = (time.perf_counter() - t_build) * 1e3
    row_ptr, col = net.host_row_ptr, net.host_col
    digest = csr_checksum(row_ptr, col)
    if digest != gold["csr_sha256"]:
        raise ParityError(f"BA({N_AGENTS}, {BA_M}) CSR checksum {digest} differs from the "
                          f"committed {gold['csr_sha256']}")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = SIRNetworkEngine(N_AGENTS, csr=net, beta=BETA, gamma=GAMMA,
                               n_seed_infected=N_SEED, seed=SEED)
        init_state = eng.state.clone()
        init_bitmaps = eng.bitmaps.clone()
        init_mark = eng.mark.clone()
        keys = eng.key_table(SIM_STEPS)
        counts = torch.zeros((SIM_STEPS, 3), dtype=torch.int32, device="cuda")
        slots = eng.slot_buffer(SIM_STEPS)
    stream.synchronize()
    sh = C.c_void_p(stream.cuda_stream)
    nbytes = C.c_size_t(N_AGENTS)
    nbytes_bm = C.c_size_t(eng.bitmaps.numel() * 4)
    nbytes_mark = C.c_size_t(eng.mark.numel() * 2)

    def enqueue_run():
        check(lib().abm_memcpy_d2d(ptr(eng.state), ptr(init_state), nbytes, sh), "reset")
        check(lib().abm_memcpy_d2d(ptr(eng.bitmaps), ptr(init_bitmaps), nbytes_bm, sh), "reset")
        check(lib().abm_memcpy_d2d(ptr(eng.mark), ptr(init_mark), nbytes_mark, sh), "reset")
        eng.launch(keys, SIM_STEPS, counts, slots=slots, t0=0, stream=stream)

    graph = None
    if not args.no_graph:
        g = C.c_void_p()
        check(lib().abm_graph_begin(sh), "graph_begin")
        enqueue_run()
        check(lib().abm_graph_end(sh, C.byref(g)), "graph_end")
        graph = g

    def one_run():
        if graph is not None:
            check(lib().abm_graph_launch(graph, sh), "graph_launch")
        else:
            enqueue_run()

    for _ in range(args.warmup):
        one_run()
    stream.synchronize()
    ref_counts = counts.cpu().numpy().copy()
    want = np.array(gold["counts"]["partitionable"], np.int32)
    if not np.array_equal(ref_counts, want):
        bad = int(np.argmax((ref_counts != want).any(axis=1)))
        raise ParityError(f"S/I/R rows differ from the oracle's from step {bad}: "
                          f"{ref_counts[bad].tolist()} vs {want[bad].tolist()}")

    # ---- timed region: K complete simulations
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    check(lib().abm_event_timer(sh, 0, None), "timer")
    for _ in range(args.steps):
        one_run()
    ms = C.c_float()
    check(lib().abm_event_timer(sh, 1, C.byref(ms)), "timer")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_s = ms.value / 1e3
    elapsed = max(wall, gpu_s)
    value = N_AGENTS * SIM_STEPS * args.steps / elapsed

    # every timed run recomputes the same simulation
    if not np.array_equal(counts.cpu().numpy(), ref_counts):
        raise ParityError("a timed replay recorded different S/I/R rows")

    # ---- step-kernel average: the event-timed graph runs hold the reset
    # copies, one slot clear and one count reduction besides the 365 step
    # kernels; all of it is charged to the step kernel (conservative).
    avg_kernel_s = gpu_s / (args.steps * SIM_STEPS)
    nnz = int(row_ptr[-1])
    algo = algorithmic_bytes(N_AGENTS, nnz)
    achieved = algo / avg_kernel_s / 1e9
    out = {
        "metric": "agent-steps/s (1M-agent SIR)",
        "value": value,
        "unit": "agent-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded Barabasi-Albert graph and seeded infections)",
        "config": {
            "workload": "network_sir_ba1M_m4_365steps",
            "n_agents": N_AGENTS, "ba_m": BA_M, "undirected_edges": nnz // 2, "csr_nnz": nnz,
            "sim_steps": SIM_STEPS, "beta": BETA, "gamma": GAMMA, "seed_infected": N_SEED,
            "rng": "jax threefry2x32, partitionable stream", "hipgraph": graph is not None,
            "graph_build": "on the device: abm_ba_edges_device + abm_csr_from_edges, "
                           "outside the timed region", "graph_build_ms": build_ms,
            "bench_step": "one full 365-step simulation",
        },
        "graph_checksum_ok": True,
        "parity": {"csr_sha256": digest, "graph_checksum_ok": True,
                   "counts_equal_oracle": True,
                   "fixture": "tests/golden/sir_ba1M_m4_365steps.json"},
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_sir_step",
            "traffic_source": None,
            "algorithmic_bytes_per_launch": algo,
            "avg_kernel_us": avg_kernel_s * 1e6,
        },
        "final_counts": ref_counts[-1].tolist(),
    }
    if graph is not None:
        lib().abm_graph_destroy(graph)
    if args.profile_run:
        return out

    # ---- data-dependent bytes: one instrumented step-by-step replay
    deg = torch.from_numpy(np.diff(row_ptr).astype(np.int64)).to("cuda")
    idx = torch.arange(N_AGENTS, device="cuda")
    with torch.cuda.stream(stream):
        eng.state.copy_(init_state)
        eng.bitmaps.copy_(init_bitmaps)
        eng.mark.copy_(init_mark)
        sl1 = eng.slot_buffer(SIM_STEPS).zero_()
        dd_bytes = []
        for t in range(SIM_STEPS):
            old = eng.state & 3
            s_mask = old == 0
            risk = s_mask & ((((eng.mark.to(torch.int32) - t) & 0xFFFF)) <= MARK_HORIZON)
            check(lib().abm_sir_step(C.byref(eng._g), ptr(eng.state), ptr(eng.mark),
                                     ptr(eng.bitmaps[t & 1]), ptr(eng.bitmaps[(t + 1) & 1]), t,
                                     ptr(keys[t]), eng.beta, eng.gamma, eng.mode,
                                     ptr(sl1[t - 1]) if t else None, ptr(sl1[t]), sh),
                  "sir_step")
            new = eng.state & 3
            marking = (new == 1) & ((old == 0) | (((t + idx) % MARK_HORIZON) == 0))
            # bytes the kernel touches (upper bound): state read, changed-state
            # writes, marks of S agents, the bitmap staged once per XCD-resident
            # copy and written once, row_ptr + col of scanned rows, and row_ptr
            # + col + 2-byte mark stores of marking rows
            dd_bytes.append(N_AGENTS + (new != old).sum() + 2 * s_mask.sum() + 8 * eng.words
                            + 8 * risk.sum() + 4 * (deg * risk).sum()
                            + 8 * marking.sum() + 6 * (deg * marking).sum())
        c1 = torch.empty((SIM_STEPS, 3), dtype=torch.int32, device="cuda")
        check(lib().abm_sir_reduce_counts(ptr(sl1), SIM_STEPS, ptr(c1), sh), "reduce")
    stream.synchronize()
    if not np.array_equal(c1.cpu().numpy(), ref_counts):
        raise ParityError("the step-by-step replay recorded different S/I/R rows")
    dd_avg = float(torch.stack(dd_bytes).double().mean().item())
    out["roofline"]["data_dependent_bytes_per_launch"] = dd_avg
    out["roofline"]["data_dependent_GBs"] = dd_avg / avg_kernel_s / 1e9
    out["roofline"]["data_dependent_frac"] = dd_avg / avg_kernel_s / 1e9 / HBM_PEAK_GBS
    traffic = measured_traffic()
    if traffic is not None:
        out["roofline"]["traffic"] = traffic["bytes_per_launch"]
        out["roofline"]["traffic_source"] = traffic["source"]
    out["roofline"]["copy_reference"] = copy_reference_gbs(torch.cuda.current_stream())
    if args.cpu_runs > 0:
        out["cpu_baseline"] = cpu_baseline(row_ptr, col, args.cpu_runs)
    return out


_SUB_DROP = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "steps", "warmup")


def sub_line(fn, args) -> dict:
    """Another program's bench line, embedded in the headline line."""
    d = fn(args)
    return {k: v for k, v in d.items() if k not in _SUB_DROP}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed full simulations")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-runs", type=int, default=5,
                    help="timed oracle runs per cpu_baseline variant (0 = skip)")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels directly instead of a hipGraph")
    ap.add_argument("--serving-requests", type=int, default=400,
                    help="requests in the serving trace (0 = skip the serving measurement)")
    ap.add_argument("--workload", default="sir",
                    choices=["sir", "deffuant", "wealth", "schelling", "generic", "serving"],
                    help="sir (default): the headline line, with the wealth, Schelling, "
                         "Deffuant and generic-model lines embedded; the others print one "
                         "program's line")
    ap.add_argument("--profile-run", action="store_true",
                    help="the SIR graph replays only (warm-up and timed), for rocprofv3: no "
                         "step-by-step replay, no other programs, no CPU baseline, no serving")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="run a 256 MiB device copy and one un-captured SIR simulation, or with "
                         "--workload deffuant one gradient evaluation (for rocprofv3 --pmc "
                         "FETCH_SIZE / WRITE_SIZE) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the engine is measured on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        return 2
    if args.pmc_probe:
        return pmc_probe(args.workload)
    import torch

    if not torch.cuda.is_available():
        print("bench.py: no ROCm GPU visible", file=sys.stderr)
        return 2
    if args.workload != "sir":
        fn = {"deffuant": bench_deffuant, "wealth": bench_wealth,
              "schelling": bench_schelling, "generic": bench_generic,
              "serving": bench_serving}[args.workload]
        print(json.dumps(fn(args)))
        return 0
    try:
        out = bench_sir(args)
    except ParityError as e:
        print(f"bench.py: parity check failed, no number reported: {e}", file=sys.stderr)
        return 3
    if not args.profile_run:
        if args.serving_requests > 0:
            out["serving"] = serving_summary(args.serving_requests)
        out["wealth"] = sub_line(bench_wealth, args)
        out["schelling"] = sub_line(bench_schelling, args)
        out["deffuant"] = sub_line(bench_deffuant, args)
        out["generic"] = sub_line(bench_generic, args)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
