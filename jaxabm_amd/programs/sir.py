"""Network SIR program on the fused HIP step kernel.

Program (written against the reference API, jaxabm/model.py:118-216):
  * one collection ``people`` whose per-agent update is the identity, and an
    env-phase callback that takes the step's ``update_key``;
  * init: ``rng, init_key = split(PRNGKey(seed))``; the first
    ``n_seed_infected`` entries of ``permutation(init_key, n)`` start infected;
  * step: ``k_inf, k_rec = split(update_key)``; ``u = uniform(k_inf, (nnz,))``,
    ``r = uniform(k_rec, (n,))``; S->I if some CSR entry e of the agent's row
    has an infected neighbour and ``u[e] < beta``; I->R if ``r[v] < gamma``;
  * record: S/I/R counts after each step (the metrics_fn, model.py:194-200).

The whole run is one device-resident launch sequence (optionally a captured
hipGraph); the host only precomputes the step-key table.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import random as jr
from .._lib import SirGraph, check, lib, ptr, require_gpu, stream_handle
from ..network import DeviceCSR, ba_csr_device, node_chunks

S, I, R = 0, 1, 2


class SIRNetworkEngine:
    """Device state and launch plumbing for one network SIR instance."""

    def __init__(self, n: int, m: int = 4, graph_seed: int = 0, beta: float = 0.05,
                 gamma: float = 0.1, n_seed_infected: int = 100, seed: int = 0,
                 mode: int | None = None,
                 csr: "tuple[np.ndarray, np.ndarray] | DeviceCSR | None" = None,
                 device: str = "cuda", chunk_cost: int | None = None):
        """``csr``: the graph as host (row_ptr, col) arrays or a DeviceCSR;
        by default BA(n, m) from ``graph_seed``, built on the device
        (network.ba_csr_device).  ``chunk_cost`` overrides the cost target of
        the agent chunks (one per block by default); a large value gives
        chunks of the maximum tile size, whatever the graph size."""
        require_gpu()
        self.n = int(n)
        self.beta = float(np.float32(beta))
        self.gamma = float(np.float32(gamma))
        self.mode = jr._mode(mode)
        self.device = torch.device(device)
        if csr is None:
            csr = ba_csr_device(self.n, m, graph_seed, self.device)
        if isinstance(csr, DeviceCSR):
            if csr.n != self.n:
                raise ValueError(f"the graph has {csr.n} nodes, the engine {self.n}")
            self.graph = csr
            row_ptr = csr.host_row_ptr   # the chunk table is cut on the host
        else:
            row_ptr, col = csr
            self.graph = DeviceCSR.from_host(row_ptr, col, self.device)
        align, node_cost, heavy_deg, heavy_cost = C.c_int32(), C.c_int64(), C.c_int32(), C.c_int64()
        per_block, blocks, max_nodes = C.c_int32(), C.c_int32(), C.c_int32()
        lib().abm_sir_chunk_params(C.byref(align), C.byref(node_cost), C.byref(heavy_deg),
                                   C.byref(heavy_cost), C.byref(per_block), C.byref(blocks),
                                   C.byref(max_nodes))
        deg = np.diff(row_ptr)
        total = (node_cost.value * self.n + int(row_ptr[-1])
                 + heavy_cost.value * int((deg > heavy_deg.value).sum()))
        target = max(1, -(-total // (blocks.value * per_block.value)))
        if chunk_cost is not None:
            target = max(1, int(chunk_cost))
        cp = node_chunks(row_ptr, align.value, node_cost.value, target, max_nodes.value,
                         heavy_deg.value, heavy_cost.value)
        self.chunks = torch.from_numpy(cp).to(self.device)
        self.n_chunks = len(cp) - 1
        self._g = SirGraph(self.n, self.graph.nnz, self.graph.row_ptr.data_ptr(),
                           self.graph.col.data_ptr(), self.chunks.data_ptr(), self.n_chunks)
        self.slots_per_step = int(lib().abm_sir_slots())
        self.state = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
        self.words = int(lib().abm_sir_bitmap_words(self.n))
        # two infected-set bitmaps; buffer 0 holds the current one between runs
        self.bitmaps = torch.zeros((2, self.words), dtype=torch.int32, device=self.device)
        self.mark = torch.zeros(self.n, dtype=torch.int16, device=self.device)
        self.t = 0
        self.n_seed_infected = int(min(n_seed_infected, self.n))
        self.seed = int(seed)
        self.rng = np.zeros(2, np.uint32)
        self.reset(seed)

    # -- init: jaxabm/model.py:118-144 with C = 1 collection
    def reset(self, seed: int | None = None) -> None:
        if seed is not None:
            self.seed = int(seed)
        keys = jr.host_split(jr.host_prng_key(self.seed), 2, self.mode)
        self.rng = keys[0].copy()
        self.t = 0
        init_key = torch.from_numpy(keys[1].view(np.int32).copy()).to(self.device)
        perm = jr.permutation(init_key, self.n, mode=self.mode)
        check(lib().abm_sir_seed(C.byref(self._g), ptr(perm), self.n_seed_infected,
                                 ptr(self.state), ptr(self.mark), ptr(self.bitmaps), self.t,
                                 stream_handle()),
              "sir_seed")

    def key_table(self, steps: int) -> torch.Tensor:
        """Per-step (k_inf, k_rec) keys for the next ``steps`` steps; advances rng."""
        tab = jr.step_key_table(self.rng, steps, 1, True, True, self.mode)
        return torch.from_numpy(np.ascontiguousarray(tab[:, 2:6]).view(np.int32)).to(self.device)

    def slot_buffer(self, steps: int) -> torch.Tensor:
        return torch.empty((max(steps, 1), self.slots_per_step, 3), dtype=torch.int32,
                           device=self.device)

    def launch(self, keys: torch.Tensor, steps: int, counts: torch.Tensor,
               slots: torch.Tensor | None = None, t0: int | None = None, stream=None) -> None:
        """Enqueue ``steps`` steps from absolute step ``t0`` (default: the
        engine's clock, which then advances) reading ``keys`` [steps, 4] and
        writing ``counts`` [steps, 3]."""
        if slots is None:
            slots = self.slot_buffer(steps)
        start = self.t if t0 is None else int(t0)
        check(lib().abm_sir_run(C.byref(self._g), ptr(self.state), ptr(self.mark),
                                ptr(self.bitmaps), start & 0xFFFFFFFF, ptr(keys), int(steps),
                                self.beta, self.gamma, self.mode, ptr(slots), ptr(counts),
                                stream_handle(stream)), "sir_run")
        if t0 is None:
            self.t += int(steps)

    def run(self, steps: int) -> np.ndarray:
        """Run ``steps`` steps; returns host counts [steps, 3] (S, I, R)."""
        keys = self.key_table(steps)
        counts = torch.empty((max(steps, 1), 3), dtype=torch.int32, device=self.device)
        self.launch(keys, steps, counts)
        return counts[:steps].cpu().numpy()

    def states(self) -> torch.Tensor:
        """Public S/I/R codes (uint8) on the device."""
        out = torch.empty_like(self.state)
        check(lib().abm_sir_states(ptr(self.state), self.n, ptr(out), stream_handle()),
              "sir_states")
        return out

    def state_host(self) -> np.ndarray:
        return self.states().cpu().numpy()
