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
    SETTLE_CHUNK = 16   # steps per launch when a run may stop once the grid settles

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
        self._peek = None
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

    def launch(self, keys, steps, counters, happy, moves, stream=None, prev=None):
        """abm_schelling_run; with ``prev`` (the counters of the step before,
        a device [2] view) abm_schelling_run_from, which continues the run the
        last call left."""
        if prev is None:
            check(lib().abm_schelling_run(C.byref(self._g), ptr(keys), int(steps), self.mode,
                                          ptr(counters), ptr(happy), ptr(moves),
                                          stream_handle(stream)), "schelling_run")
        else:
            check(lib().abm_schelling_run_from(C.byref(self._g), ptr(keys), int(steps),
                                               self.mode, ptr(prev), ptr(counters), ptr(happy),
                                               ptr(moves), stream_handle(stream)),
                  "schelling_run_from")

    def run(self, steps: int, chunk: int = SETTLE_CHUNK):
        """Run ``steps`` steps; returns host (fraction happy, moves) arrays.
        ``chunk``: steps per launch, with the settled-grid exit of
        run_settling (0: all steps in one launch)."""
        keys = self.key_table(steps)
        counters = torch.empty((max(steps, 1), 2), dtype=torch.int32, device=self.device)
        happy = torch.empty(max(steps, 1), dtype=torch.float32, device=self.device)
        moves = torch.empty(max(steps, 1), dtype=torch.int32, device=self.device)
        self.run_settling(keys, steps, counters, happy, moves, chunk)
        return happy[:steps].cpu().numpy(), moves[:steps].cpu().numpy()

    def run_settling(self, keys, steps, counters, happy, moves, chunk=SETTLE_CHUNK, stream=None,
                     launch=None) -> int:
        """Launch the steps ``chunk`` at a time and stop once the grid settles.

        After each launch the happy count of its last step is copied back
        (pinned memory, an event); the host reads launch i's count before it
        queues launch i + 2.  A step whose previous step found every agent
        happy is the identity, and so is every step after it: once launch i
        ends with all agents happy, launch i + 1 (already queued) runs identity
        steps and the records of the steps after it are written by one
        abm_schelling_settled call instead, the values those steps would
        record.  ``launch(i, s0, s1)`` queues steps s0..s1-1 as launch i
        (default on this stream: abm_schelling_run for the first,
        abm_schelling_run_from for the others).  Returns the number of steps
        launched."""
        st = stream or torch.cuda.current_stream(self.device)
        if launch is None:
            def launch(i, s0, s1):
                self.launch(keys[s0:], s1 - s0, counters[s0:], happy[s0:], moves[s0:], st,
                            prev=counters[s0 - 1] if s0 else None)
        chunk = chunk if chunk > 0 else max(steps, 1)
        bounds = [(s0, min(s0 + chunk, steps)) for s0 in range(0, steps, chunk)]
        if len(bounds) <= 2:   # nothing to stop early
            for i, (s0, s1) in enumerate(bounds):
                launch(i, s0, s1)
            return steps
        if self._peek is None or self._peek.numel() < len(bounds):
            self._peek = torch.empty(len(bounds), dtype=torch.int32, pin_memory=True)
        events = []
        with torch.cuda.stream(st):
            for i, (s0, s1) in enumerate(bounds):
                if i >= 2:
                    events[i - 2].synchronize()
                    if int(self._peek[i - 2]) == self.n_agents:
                        check(lib().abm_schelling_settled(
                            C.byref(self._g), ptr(counters[s0 - 1]), steps - s0, ptr(counters[s0:]),
                            ptr(happy[s0:]), ptr(moves[s0:]), stream_handle(st)),
                            "schelling_settled")
                        return s0
                launch(i, s0, s1)
                self._peek[i:i + 1].copy_(counters[s1 - 1, 0:1], non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(st)
                events.append(ev)
        return steps

    def grid(self) -> np.ndarray:
        return self.cells.cpu().numpy().reshape(self.H, self.W)

    def positions(self) -> np.ndarray:
        return self.pos[: self.n_agents].cpu().numpy()
