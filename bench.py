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
    byte per CSR entry) over the kernel's average duration measured with HIP
    events on the launch stream, against MI355X's 8 TB/s;
  * the data-dependent bytes the early-exit kernel actually has to touch;
  * cpu_baseline: the C oracle (a restatement of the reference semantics, one
    thread) running the same full simulation on the host.
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
    import oracle as O

    times = []
    for _ in range(runs):
        st, rng = O.sir_init(N_AGENTS, N_SEED, SEED, O.MODE_PARTITIONABLE)
        t0 = time.perf_counter()
        O.sir_run(row_ptr, col, st, rng, BETA, GAMMA, SIM_STEPS, O.MODE_PARTITIONABLE)
        times.append(time.perf_counter() - t0)
    best = min(times)
    return {
        "value": N_AGENTS * SIM_STEPS / best,
        "unit": "agent-steps/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"full 365-step 1M-agent SIR simulation, C oracle (restatement of the "
                   f"reference semantics, not JAX), single thread, best of {runs}; "
                   f"host CPU: {cpu_model()}, nproc={os.cpu_count()}"),
        "run_s": best,
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed full simulations")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-runs", type=int, default=2, help="oracle runs for cpu_baseline (0 = skip)")
    ap.add_argument("--no-graph", action="store_true", help="launch kernels directly instead of a hipGraph")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding is a design note, not a measured path.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from jaxabm_amd._lib import check, lib, ptr
    from jaxabm_amd.network import ba_csr
    from jaxabm_amd.programs.sir import SIRNetworkEngine

    if not torch.cuda.is_available():
        print("bench.py: no ROCm GPU visible", file=sys.stderr)
        return 2

    row_ptr, col = ba_csr(N_AGENTS, BA_M, GRAPH_SEED)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        eng = SIRNetworkEngine(N_AGENTS, csr=(row_ptr, col), beta=BETA, gamma=GAMMA,
                               n_seed_infected=N_SEED, seed=SEED)
        init_state = eng.state.clone()
        rng0 = eng.rng.copy()
        keys = eng.key_table(SIM_STEPS)
        counts = torch.zeros((SIM_STEPS, 3), dtype=torch.int32, device="cuda")
    stream.synchronize()
    sh = C.c_void_p(stream.cuda_stream)
    nbytes = C.c_size_t(N_AGENTS)

    def enqueue_run():
        check(lib().abm_memcpy_d2d(ptr(eng.state), ptr(init_state), nbytes, sh), "reset")
        eng.launch(keys, SIM_STEPS, counts, stream=stream)

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
    assert np.array_equal(counts.cpu().numpy(), ref_counts)
    assert (ref_counts.sum(axis=1) == N_AGENTS).all()

    # ---- per-kernel timing: one instrumented run, events around each step kernel
    nnz = int(row_ptr[-1])
    deg = torch.from_numpy(np.diff(row_ptr).astype(np.int64)).to("cuda")
    with torch.cuda.stream(stream):
        eng.state.copy_(init_state)
        c1 = torch.zeros((SIM_STEPS, 3), dtype=torch.int32, device="cuda")
        bufs = [eng.state, eng.scratch]
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(SIM_STEPS)]
        dd_bytes = []
        for t in range(SIM_STEPS):
            cur = bufs[t & 1]
            s_mask = cur == 0
            dd_bytes.append(2 * N_AGENTS + 8 * s_mask.sum() + 5 * (deg * s_mask).sum())
            evs[t][0].record(stream)
            check(lib().abm_sir_step(C.byref(eng._g), ptr(cur), ptr(bufs[(t + 1) & 1]),
                                     ptr(keys[t]), eng.beta, eng.gamma, eng.mode, ptr(c1[t]), sh),
                  "sir_step")
            evs[t][1].record(stream)
    stream.synchronize()
    kern_ms = [a.elapsed_time(b) for a, b in evs]
    assert np.array_equal(c1.cpu().numpy(), ref_counts)
    avg_kernel_s = sum(kern_ms) / len(kern_ms) / 1e3
    algo = algorithmic_bytes(N_AGENTS, nnz)
    achieved = algo / avg_kernel_s / 1e9
    dd_avg = float(torch.stack(dd_bytes).double().mean().item())

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
            "bench_step": "one full 365-step simulation",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "k_sir_step",
            "algorithmic_bytes_per_launch": algo,
            "avg_kernel_us": avg_kernel_s * 1e6,
            "data_dependent_bytes_per_launch": dd_avg,
            "data_dependent_GBs": dd_avg / avg_kernel_s / 1e9,
        },
        "final_counts": ref_counts[-1].tolist(),
    }
    if args.cpu_runs > 0:
        out["cpu_baseline"] = cpu_baseline(row_ptr, col, args.cpu_runs)
    if graph is not None:
        lib().abm_graph_destroy(graph)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
