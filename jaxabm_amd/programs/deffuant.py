"""Differentiable Deffuant opinion model: a batch of (eps, mu) instances on
the HIP step kernels, with the hand-written adjoint behind
``torch.autograd.Function``.

Program (written against the reference API, jaxabm/model.py:118-216; loss and
gradient set-up after jaxabm/analysis.py:489-531, where the reference takes
``jit(grad(loss))`` through the unrolled run):
  * init: rng, init_key = split(PRNGKey(seed)); agent keys = split(init_key, n);
    x_i = uniform(agent_key_i), the same for every instance;
  * step t: the env-phase key uk_t of the reference schedule; perm =
    permutation(uk_t, n); pairs (perm[2i], perm[2i+1]);
    d = x_c - x_a; delta = mu * sigmoid((eps - |d|) / 0.02) * d;
    x_a += delta; x_c -= delta;
  * loss per instance: squared error of [mean, variance, 32-bin soft
    histogram] of the final opinions against a target (float64).

All instances share the seed, hence the pairings (common random numbers):
the pairings of every step are drawn once per model by one batched sort.
The forward keeps a checkpoint every ``ckpt_every`` steps; the backward
recomputes each segment from its checkpoint and sweeps it in reverse.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .. import random as jr
from .._lib import (DEFFUANT_TAPE_LAST, DeffuantProblem, check, lib, ptr, require_gpu,
                    stream_handle)

STATS_LEN = 34
HIST_BINS = 32


def _default_pair_blocks(n: int, batch: int) -> int:
    """Workgroup rows of the step grids: about 4096 wavefronts in all (16 per
    CU on 256 CUs), at least one pair per wavefront."""
    col_blocks = max(1, -(-batch // 256))
    rows = max(1, 1024 // col_blocks)
    return int(max(1, min(rows, (n // 2 + 3) // 4, 65535)))


def _default_stat_chunks(n: int, batch: int) -> int:
    col_blocks = max(1, -(-batch // 256))
    return int(max(1, min(n, 1024 // col_blocks, 65535)))


class DeffuantModel:
    """Device state of a batch of Deffuant instances sharing one seed."""

    def __init__(self, n: int = 65536, batch: int = 256, steps: int = 200, seed: int = 0,
                 ckpt_every: int = 20, mode: int | None = None, pair_blocks: int | None = None,
                 stat_chunks: int | None = None, device: str = "cuda"):
        require_gpu()
        if n < 2 or batch < 1 or steps < 0 or ckpt_every < 1:
            raise ValueError("need n >= 2, batch >= 1, steps >= 0, ckpt_every >= 1")
        self.n, self.batch, self.steps = int(n), int(batch), int(steps)
        self.ckpt_every = int(ckpt_every)
        self.mode = jr._mode(mode)
        self.device = torch.device(device)
        self.pair_blocks = int(pair_blocks or _default_pair_blocks(self.n, self.batch))
        self.stat_chunks = int(stat_chunks or _default_stat_chunks(self.n, self.batch))
        dev, n, B = self.device, self.n, self.batch
        keys = jr.host_split(jr.host_prng_key(seed), 2, self.mode)
        self.rng0 = keys[0].copy()
        # env-phase keys of every step, then all pairings in one batched sort
        rng = self.rng0.copy()
        tab = jr.step_key_table(rng, self.steps, 1, True, False, self.mode)
        self.rng_end = rng
        self.perms = torch.empty((max(self.steps, 1), n), dtype=torch.int32, device=dev)
        if self.steps:
            uk = torch.from_numpy(np.ascontiguousarray(tab[:, 2:4]).view(np.int32)).to(dev)
            self.perms[: self.steps] = jr.permutation_each(uk, n, self.mode)
        n_ckpt = max(1, -(-self.steps // self.ckpt_every))
        self.ckpt = torch.empty((n_ckpt, n, B), dtype=torch.float32, device=dev)
        ik = torch.from_numpy(keys[1].view(np.int32).copy()).to(dev)
        check(lib().abm_deffuant_init(ptr(ik), n, B, ptr(self.ckpt[0]), self.mode,
                                      stream_handle()), "deffuant_init")
        self.work = torch.empty((n, B), dtype=torch.float32, device=dev)
        self.traj = torch.empty((max(self.ckpt_every - 1, 1), n, B), dtype=torch.float32,
                                device=dev)
        self.lam = torch.empty((n, B), dtype=torch.float32, device=dev)
        f64 = torch.float64
        self.acc = torch.empty((self.pair_blocks, 2, B), dtype=f64, device=dev)
        self.partial = torch.empty((self.stat_chunks, STATS_LEN, B), dtype=f64, device=dev)
        self.sums = torch.empty((STATS_LEN, B), dtype=f64, device=dev)
        self.coef = torch.empty((STATS_LEN, B), dtype=f64, device=dev)
        self.stats = torch.empty((B, STATS_LEN), dtype=f64, device=dev)
        self.loss = torch.empty((B,), dtype=f64, device=dev)
        self.grads = torch.empty((2, B), dtype=f64, device=dev)
        self.target = torch.zeros((STATS_LEN,), dtype=f64, device=dev)
        self.weights = torch.ones((STATS_LEN,), dtype=f64, device=dev)
        self.eps = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.mu = torch.zeros((B,), dtype=torch.float32, device=dev)
        self.version = 0
        self._backward_ready = False
        self._tape_fresh = False   # traj holds the last segment from the forward

    @property
    def x0(self) -> torch.Tensor:
        return self.ckpt[0]

    def final_state(self) -> torch.Tensor:
        return self.work if self.steps else self.ckpt[0]

    def _problem(self, flags: int = 0) -> DeffuantProblem:
        p = DeffuantProblem()
        p.n, p.batch, p.steps, p.flags = self.n, self.batch, self.steps, flags
        p.ckpt_every, p.pair_blocks, p.stat_chunks = (self.ckpt_every, self.pair_blocks,
                                                      self.stat_chunks)
        for name in ("perms", "eps", "mu", "work", "ckpt", "traj", "lam", "acc", "partial",
                     "sums", "coef", "target", "stats", "loss", "grads", "weights"):
            setattr(p, name, getattr(self, name).data_ptr())
        return p

    def forward(self, eps, mu, target, weights=None) -> torch.Tensor:
        """Run the batch; returns the per-instance float64 losses (device).
        ``weights`` (34 values, default all 1) weight each statistic's squared
        error."""
        eps = torch.as_tensor(eps, dtype=torch.float32, device=self.device).reshape(-1)
        mu = torch.as_tensor(mu, dtype=torch.float32, device=self.device).reshape(-1)
        if eps.numel() != self.batch or mu.numel() != self.batch:
            raise ValueError(f"eps and mu need {self.batch} values")
        self.eps.copy_(eps)
        self.mu.copy_(mu)
        self.target.copy_(torch.as_tensor(np.asarray(target, np.float64).reshape(STATS_LEN)))
        w = np.ones(STATS_LEN) if weights is None else np.asarray(weights, np.float64)
        self.weights.copy_(torch.as_tensor(w.reshape(STATS_LEN)))
        check(lib().abm_deffuant_forward(C.byref(self._problem(DEFFUANT_TAPE_LAST)),
                                         stream_handle()), "deffuant_forward")
        self.version += 1
        self._backward_ready = True
        self._tape_fresh = True
        return self.loss.clone()

    def backward(self) -> tuple[torch.Tensor, torch.Tensor]:
        """dL_b/deps_b and dL_b/dmu_b (float64, device) of the last forward."""
        if not self._backward_ready:
            raise RuntimeError("backward needs a forward on this model first")
        flags = DEFFUANT_TAPE_LAST if self._tape_fresh else 0
        self._tape_fresh = False   # this sweep overwrites traj
        check(lib().abm_deffuant_backward(C.byref(self._problem(flags)), stream_handle()),
              "deffuant_backward")
        return self.grads[0].clone(), self.grads[1].clone()


def target_stats(n: int = 65536, steps: int = 200, seed: int = 0, eps: float = 0.2,
                 mu: float = 0.3, mode: int | None = None) -> np.ndarray:
    """Statistics of one run at (eps, mu): the calibration target."""
    m = DeffuantModel(n, 1, steps, seed, ckpt_every=max(steps, 1), mode=mode)
    m.forward([eps], [mu], np.zeros(STATS_LEN))
    return m.stats[0].cpu().numpy()


class DeffuantLoss(torch.autograd.Function):
    """loss[b] = L(eps[b], mu[b]) through ``model``; backward runs the
    hand-written adjoint of the same run."""

    @staticmethod
    def forward(ctx, eps, mu, model: DeffuantModel, target):
        loss = model.forward(eps.detach(), mu.detach(), target)
        ctx.model = model
        ctx.version = model.version
        ctx.dtypes = (eps.dtype, mu.dtype)
        return loss

    @staticmethod
    def backward(ctx, grad_loss):
        model = ctx.model
        if model.version != ctx.version:
            raise RuntimeError("the model ran another forward since this loss was computed")
        ge, gm = model.backward()
        g = grad_loss.to(torch.float64)
        return (g * ge).to(ctx.dtypes[0]), (g * gm).to(ctx.dtypes[1]), None, None


def deffuant_loss(eps: torch.Tensor, mu: torch.Tensor, model: DeffuantModel, target):
    return DeffuantLoss.apply(eps, mu, model, target)
