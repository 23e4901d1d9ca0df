"""Wealth-exchange random walk on the HIP engine (one workgroup per instance,
whole run in one launch).

Program (written against the reference API, jaxabm/model.py:118-216,
jaxabm/agent.py:92-177; Gini after examples/models/advanced_economic_model.py:
1798-1811):
  * one collection ``walkers`` (n agents) and an env-phase callback;
  * init: agent keys = split(init_key, n); kx, ky = split(agent key);
    x = uniform(kx), y = uniform(ky), wealth = 1;
  * agent phase: k1, k2 = split(agent key); x = mod(x + 0.01 normal(k1), 1),
    same for y with k2;
  * env phase: t = randint(update_key, (n,), 0, n-1); t += (t >= i); every
    agent with wealth > 0 gives one unit to agent t (exact scatter-add);
  * record: mean(wealth) and Gini(wealth) each step.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import random as jr
from .._lib import (WEALTH_GRID_ACC_WORDS, WEALTH_INST, check, lib, ptr, require_gpu,
                    stream_handle)


# above the LDS capacity of one workgroup (8,000 agents), a run spreads over
# the whole GPU
GRID_MIN_AGENTS = 8000


class WealthEngine:
    """Device state of one wealth-exchange instance."""

    def __init__(self, n: int = 1000, seed: int = 0, mode: int | None = None,
                 device: str = "cuda"):
        require_gpu()
        if n < 2:
            raise ValueError("the wealth exchange needs at least 2 agents")
        self.n = int(n)
        self.mode = jr._mode(mode)
        self.device = torch.device(device)
        self.x = torch.empty(self.n, dtype=torch.float32, device=self.device)
        self.y = torch.empty(self.n, dtype=torch.float32, device=self.device)
        self.w = torch.empty(self.n, dtype=torch.int32, device=self.device)
        self.inc = torch.zeros(self.n, dtype=torch.int32, device=self.device)
        self.hist = torch.zeros(self.n + 1, dtype=torch.int32, device=self.device)
        self.rng = np.zeros(2, np.uint32)
        self.reset(seed)

    def reset(self, seed: int) -> None:
        """jaxabm/model.py:118-144 with one collection: rng, init_key = split(PRNGKey(seed))."""
        keys = jr.host_split(jr.host_prng_key(seed), 2, self.mode)
        self.rng = keys[0].copy()
        ik = torch.from_numpy(keys[1].view(np.int32).copy()).to(self.device)
        check(lib().abm_wealth_init(ptr(ik), self.n, ptr(self.x), ptr(self.y), ptr(self.w),
                                    self.mode, stream_handle()), "wealth_init")

    def key_table(self, steps: int) -> torch.Tensor:
        tab = jr.step_key_table(self.rng, steps, 1, True, False, self.mode)
        return torch.from_numpy(np.ascontiguousarray(tab).view(np.int32)).to(self.device)

    def draw_buffer(self, steps: int) -> torch.Tensor:
        """Scratch for the run's draws: [steps][3][n] 32-bit words."""
        return torch.empty(max(steps, 1) * 3 * self.n, dtype=torch.int32, device=self.device)

    def descriptor(self, keys: torch.Tensor, steps: int, mean: torch.Tensor,
                   gini: torch.Tensor, draws: torch.Tensor | None = None) -> np.ndarray:
        d = np.zeros(1, WEALTH_INST)
        d["n"] = self.n
        d["x"], d["y"], d["w"] = self.x.data_ptr(), self.y.data_ptr(), self.w.data_ptr()
        d["inc"], d["hist"] = self.inc.data_ptr(), self.hist.data_ptr()
        d["keys"], d["steps"] = keys.data_ptr(), steps
        d["mean_out"], d["gini_out"] = mean.data_ptr(), gini.data_ptr()
        d["draws"] = draws.data_ptr() if draws is not None else 0
        return d

    def run(self, steps: int):
        """Run ``steps`` steps; returns host (mean, gini) float32 arrays.
        Up to GRID_MIN_AGENTS agents one workgroup runs the whole simulation
        out of LDS; larger instances spread every step over the whole GPU
        (abm_wealth_run_grid).  Same results either way."""
        keys = self.key_table(steps)
        mean = torch.empty(max(steps, 1), dtype=torch.float32, device=self.device)
        gini = torch.empty(max(steps, 1), dtype=torch.float32, device=self.device)
        if self.n > GRID_MIN_AGENTS:
            # draws computed inline by the step kernels: no draw buffer
            host = self.descriptor(keys, steps, mean, gini)
            acc = torch.empty(WEALTH_GRID_ACC_WORDS, dtype=torch.int32, device=self.device)
            check(lib().abm_wealth_run_grid(host.ctypes.data, ptr(acc), self.mode,
                                            stream_handle()), "wealth_run_grid")
        else:
            draws = self.draw_buffer(steps)   # held until the run has read it
            host = self.descriptor(keys, steps, mean, gini, draws)
            desc = torch.from_numpy(host.view(np.uint8)).to(self.device)
            check(lib().abm_wealth_run_batch(ptr(desc), 1, self.n, self.mode, stream_handle()),
                  "wealth_run")
        return mean[:steps].cpu().numpy(), gini[:steps].cpu().numpy()


def run_batch(engines: list[WealthEngine], steps: list[int], stream=None):
    """Run several instances in one launch (one workgroup each); returns a list
    of (mean, gini) host arrays."""
    if not engines:
        return []
    mode = engines[0].mode
    if any(e.mode != mode for e in engines):
        raise ValueError("a batch shares one RNG stream mode")
    dev = engines[0].device
    keep, descs, outs = [], [], []
    for e, s in zip(engines, steps):
        keys = e.key_table(s)
        mean = torch.empty(max(s, 1), dtype=torch.float32, device=dev)
        gini = torch.empty(max(s, 1), dtype=torch.float32, device=dev)
        draws = e.draw_buffer(s)
        keep += [keys, mean, gini, draws]
        descs.append(e.descriptor(keys, s, mean, gini, draws))
        outs.append((mean, gini, s))
    table = torch.from_numpy(np.concatenate(descs).view(np.uint8)).to(dev)
    max_n = max(e.n for e in engines)
    check(lib().abm_wealth_run_batch(ptr(table), len(engines), max_n, mode, stream_handle(stream)),
          "wealth_run_batch")
    return [(m[:s].cpu().numpy(), g[:s].cpu().numpy()) for m, g, s in outs]
