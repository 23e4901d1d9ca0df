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
from ..network import DeviceCSR, ba_csr, heavy_rows

S, I, R = 0, 1, 2


class SIRNetworkEngine:
    """Device state and launch plumbing for one network SIR instance."""

    def __init__(self, n: int, m: int = 4, graph_seed: int = 0, beta: float = 0.05,
                 gamma: float = 0.1, n_seed_infected: int = 100, seed: int = 0,
                 mode: int | None = None, csr: tuple[np.ndarray, np.ndarray] | None = None,
                 device: str = "cuda"):
        require_gpu()
        self.n = int(n)
        self.beta = float(np.float32(beta))
        self.gamma = float(np.float32(gamma))
        self.mode = jr._mode(mode)
        self.device = torch.device(device)
        if csr is None:
            csr = ba_csr(self.n, m, graph_seed)
        row_ptr, col = csr
        self.graph = DeviceCSR.from_host(row_ptr, col, self.device)
        hv = heavy_rows(row_ptr, int(lib().abm_sir_heavy_threshold()))
        self.heavy = torch.from_numpy(hv if len(hv) else np.zeros(1, np.int32)).to(self.device)
        self.n_heavy = len(hv)
        self._g = SirGraph(self.n, self.graph.nnz, self.graph.row_ptr.data_ptr(),
                           self.graph.col.data_ptr(), self.heavy.data_ptr(), self.n_heavy)
        self.state = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
        self.scratch = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
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
        init_key = torch.from_numpy(keys[1].view(np.int32).copy()).to(self.device)
        perm = jr.permutation(init_key, self.n, mode=self.mode)
        check(lib().abm_sir_seed(self.n, ptr(perm), self.n_seed_infected, ptr(self.state),
                                 stream_handle()), "sir_seed")

    def key_table(self, steps: int) -> torch.Tensor:
        """Per-step (k_inf, k_rec) keys for the next ``steps`` steps; advances rng."""
        tab = jr.step_key_table(self.rng, steps, 1, True, True, self.mode)
        return torch.from_numpy(np.ascontiguousarray(tab[:, 2:6]).view(np.int32)).to(self.device)

    def launch(self, keys: torch.Tensor, steps: int, counts: torch.Tensor, stream=None) -> None:
        """Enqueue ``steps`` steps reading ``keys`` [steps, 4] and writing ``counts`` [steps, 3]."""
        check(lib().abm_sir_run(C.byref(self._g), ptr(self.state), ptr(self.scratch), ptr(keys),
                                int(steps), self.beta, self.gamma, self.mode, ptr(counts),
                                stream_handle(stream)), "sir_run")

    def run(self, steps: int) -> np.ndarray:
        """Run ``steps`` steps; returns host counts [steps, 3] (S, I, R)."""
        keys = self.key_table(steps)
        counts = torch.empty((max(steps, 1), 3), dtype=torch.int32, device=self.device)
        self.launch(keys, steps, counts)
        return counts[:steps].cpu().numpy()

    def state_host(self) -> np.ndarray:
        return self.state.cpu().numpy()
