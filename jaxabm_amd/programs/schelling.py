"""Schelling segregation on a periodic Grid with the HIP step kernels.

Program (the reference's Grid stores only shape and the periodic flag,
jaxabm/agentpy.py:465-527; step schedule jaxabm/model.py:146-216):
  * init: u = uniform(PRNGKey(seed), (H*W,)); type 1 if u < 0.45, 2 if
    u < 0.9, else empty; agent k is the k-th occupied cell in row-major order;
    rng, init_key = split(PRNGKey(seed)) (one collection);
  * agent phase (agent keys = split(ck, n_agents)): Moore-8 on the torus,
    happy iff same/occupied >= 0.375 (isolated agents are happy); an unhappy
    agent draws j = randint(agent_key, (), 0, n_empty) and claims the j-th
    empty cell (row-major) of the pre-step grid;
  * env phase: each claimed cell goes to its lowest-id claimant; winners move;
  * record: fraction happy (agent phase) and number of moves.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import random as jr
from .._lib import SchellingGrid, check, lib, ptr, require_gpu, stream_handle


class SchellingEngine:
    def __init__(self, height: int = 512, width: int = 512, seed: int = 0,
                 mode: int | None = None, device: str = "cuda"):
        require_gpu()
        self.H, self.W = int(height), int(width)
        self.mode = jr._mode(mode)
        self.device = torch.device(device)
        nc = self.H * self.W
        dev = self.device
        self.cells = torch.zeros(nc, dtype=torch.int8, device=dev)
        self.id = torch.empty(nc, dtype=torch.int32, device=dev)
        self.claim = torch.empty(nc, dtype=torch.int32, device=dev)
        self.pos = torch.empty(nc, dtype=torch.int32, device=dev)
        self.type = torch.empty(nc, dtype=torch.int8, device=dev)
        self.target = torch.empty(nc, dtype=torch.int32, device=dev)
        self.empty = torch.empty(nc, dtype=torch.int32, device=dev)
        self.block_counts = torch.empty(int(lib().abm_schelling_block_count(nc)),
                                        dtype=torch.int32, device=dev)
        self._g = SchellingGrid(self.H, self.W, 0, self.cells.data_ptr(), self.id.data_ptr(),
                                self.claim.data_ptr(), self.pos.data_ptr(), self.type.data_ptr(),
                                self.target.data_ptr(), self.empty.data_ptr(),
                                self.block_counts.data_ptr())
        self.n_agents = 0
        self.rng = np.zeros(2, np.uint32)
        self.reset(seed)

    def reset(self, seed: int) -> None:
        key = torch.from_numpy(jr.host_prng_key(seed).view(np.int32).copy()).to(self.device)
        na = C.c_int32()
        check(lib().abm_schelling_init(C.byref(self._g), ptr(key), self.mode, C.byref(na),
                                       stream_handle()), "schelling_init")
        self.n_agents = int(na.value)
        self.rng = jr.host_split(jr.host_prng_key(seed), 2, self.mode)[0].copy()

    def key_table(self, steps: int) -> torch.Tensor:
        tab = jr.step_key_table(self.rng, steps, 1, True, False, self.mode)
        ck = np.ascontiguousarray(tab[:, 0:2])
        return torch.from_numpy(ck.view(np.int32)).to(self.device)

    def launch(self, keys, steps, counters, happy, moves, stream=None):
        check(lib().abm_schelling_run(C.byref(self._g), ptr(keys), int(steps), self.mode,
                                      ptr(counters), ptr(happy), ptr(moves),
                                      stream_handle(stream)), "schelling_run")

    def run(self, steps: int):
        """Run ``steps`` steps; returns host (fraction happy, moves) arrays."""
        keys = self.key_table(steps)
        counters = torch.empty((max(steps, 1), 2), dtype=torch.int32, device=self.device)
        happy = torch.empty(max(steps, 1), dtype=torch.float32, device=self.device)
        moves = torch.empty(max(steps, 1), dtype=torch.int32, device=self.device)
        self.launch(keys, steps, counters, happy, moves)
        return happy[:steps].cpu().numpy(), moves[:steps].cpu().numpy()

    def grid(self) -> np.ndarray:
        return self.cells.cpu().numpy().reshape(self.H, self.W)

    def positions(self) -> np.ndarray:
        return self.pos[: self.n_agents].cpu().numpy()
