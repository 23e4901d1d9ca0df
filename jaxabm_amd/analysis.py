"""Sensitivity analysis and calibration (jaxabm/analysis.py:23-2754).

* :class:`SensitivityAnalysis` -- Latin hypercube samples drawn from the
  threefry stream as the reference draws them (jaxabm/analysis.py:67-95:
  key, sub = split(key); points = linspace(0, 1, n+1)[:-1] + uniform(sub,
  (n,)) / n; per parameter key, sub = split(key) and permutation(sub,
  points)), one model run per sample with seed i + 1000
  (jaxabm/analysis.py:97-165), correlation-based indices
  (jaxabm/analysis.py:167-203).
* :class:`ModelCalibrator` -- loss types and robust evaluation as in
  jaxabm/analysis.py:401-487; optimisers "adam"/"sgd" (gradient of the loss
  at seed 42, jaxabm/analysis.py:489-531 and 1309-1394), "es", "pso" and
  "cem".  The reference differentiates the whole simulation with
  ``jax.grad``; here a model that provides ``loss_gradient`` (the Deffuant
  program's hand-written adjoint) supplies the exact gradient, a generic
  model torch autograd through its eager run (parameters as float32
  tensors), and a model whose run reads a parameter on the host a central
  finite difference of its runs.  The reinforcement-learning
  calibrators of the reference are not part of this engine.
"""
from __future__ import annotations

from typing import Any, Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import random as jr
from .core import ModelConfig

ModelFactory = Callable[..., Any]


def _last(results, metric):
    data = results._data if hasattr(results, "_data") else results
    if metric not in data or data[metric] is None:
        return None
    v = data[metric]
    if isinstance(v, torch.Tensor):
        v = v.detach().cpu().numpy()
    if hasattr(v, "__len__") and not isinstance(v, str):
        v = v[-1]
    return float(v)


class _Concrete(Exception):
    """A host read of a tensor that carries a gradient."""


class _NoConcretization(torch.overrides.TorchFunctionMode):
    """Raises on a host read (float(), int(), bool(), .item(), .tolist(),
    .numpy()) of a tensor that requires grad, as JAX refuses to concretize a
    traced value under jax.grad: a value read on the host is cut from the
    gradient, so the autograd gradient would be silently partial."""

    _READS = frozenset({torch.Tensor.__float__, torch.Tensor.__int__, torch.Tensor.__bool__,
                        torch.Tensor.__index__, torch.Tensor.item, torch.Tensor.tolist,
                        torch.Tensor.numpy})

    def __torch_function__(self, func, types, args=(), kwargs=None):
        if func in self._READS and args and isinstance(args[0], torch.Tensor) and \
                args[0].requires_grad:
            raise _Concrete(getattr(func, "__name__", str(func)))
        return func(*args, **(kwargs or {}))


def _servable(model) -> bool:
    from .programs.library import servable

    return servable(model)


def _run_models(models, steps=None):
    from .programs.library import run_models

    return run_models(models, steps)


def _pyplot():
    try:
        import matplotlib.pyplot as plt
    except ImportError as e:
        raise ImportError("plotting needs matplotlib, which is not installed") from e
    return plt


class SensitivityAnalysis:
    def __init__(self, model_factory: ModelFactory, param_ranges: Dict[str, Tuple[float, float]],
                 metrics_of_interest: List[str], num_samples: int = 100, seed: int = 0):
        self.model_factory = model_factory
        self.param_ranges = param_ranges
        self.metrics_of_interest = metrics_of_interest
        self.num_samples = num_samples
        self.key = jr.host_prng_key(seed)
        self.samples = self._generate_lhs_samples()
        self.results = None

    def _generate_lhs_samples(self) -> np.ndarray:
        n = self.num_samples
        self.key, sub = jr.host_split(self.key, 2)
        points = np.linspace(0.0, 1.0, n + 1, dtype=np.float32)[:-1]
        jitter = jr.uniform(sub, (n,)).cpu().numpy()
        points = (points + jitter / np.float32(n)).astype(np.float32)
        pts = torch.from_numpy(points).to(jr._device())
        samples = np.zeros((n, len(self.param_ranges)), np.float32)
        for i, _ in enumerate(self.param_ranges):
            self.key, sub = jr.host_split(self.key, 2)
            samples[:, i] = jr.permutation(sub, pts).cpu().numpy()
        for i, (_, (lo, hi)) in enumerate(self.param_ranges.items()):
            samples[:, i] = samples[:, i] * np.float32(hi - lo) + np.float32(lo)
        return samples

    def run(self, verbose: bool = True) -> Dict[str, np.ndarray]:
        """One run per sample with seed 1000 + i.  Library models with a
        server request (wealth, network SIR) run as one continuous batch,
        with the results of running them one after another."""
        if verbose:
            print(f"Running sensitivity analysis with {self.num_samples} samples...")
        names = list(self.param_ranges.keys())
        out = {m: np.zeros(self.num_samples, np.float32) for m in self.metrics_of_interest}
        sample_params = [{p: float(self.samples[i, j]) for j, p in enumerate(names)}
                         for i in range(self.num_samples)]
        batch = None
        if self.num_samples > 1:
            first = self.model_factory(params=sample_params[0], config=ModelConfig(seed=1000))
            if _servable(first):
                models = [first] + [self.model_factory(params=sample_params[i],
                                                       config=ModelConfig(seed=i + 1000))
                                    for i in range(1, self.num_samples)]
                batch = _run_models(models)
        for i in range(self.num_samples):
            params = sample_params[i]
            if verbose:
                print(f"\nSample {i+1}/{self.num_samples}")
                print("Parameters: " + ", ".join(f"{k}={v:.4f}" for k, v in params.items()))
            if batch is not None:
                results = batch[i]
            else:
                model = self.model_factory(params=params, config=ModelConfig(seed=i + 1000))
                results = model.run()
            for m in self.metrics_of_interest:
                v = _last(results, m)
                if v is not None:
                    out[m][i] = v
                    if verbose:
                        print(f"  {m}: {v:.4f}")
        self.results = out
        return out

    def sobol_indices(self) -> Dict[str, Dict[str, float]]:
        """Squared correlation of each parameter with each metric (a proxy of
        first-order indices, as the reference computes it)."""
        if self.results is None:
            raise ValueError("Must run sensitivity analysis before calculating indices")
        names = list(self.param_ranges.keys())
        indices = {}
        for metric, values in self.results.items():
            v = (values - values.mean()) / (values.std() + 1e-8)
            indices[metric] = {}
            for i, p in enumerate(names):
                x = self.samples[:, i]
                xn = (x - x.mean()) / (x.std() + 1e-8)
                indices[metric][p] = float(np.mean(xn * v) ** 2)
        return indices

    def morris_indices(self) -> Dict[str, Dict[str, float]]:
        """Mean absolute elementary effect estimated from the samples: |slope|
        of a least-squares line of the metric on each parameter, times the
        parameter range."""
        if self.results is None:
            raise ValueError("Must run sensitivity analysis before calculating indices")
        names = list(self.param_ranges.keys())
        out = {}
        for metric, values in self.results.items():
            out[metric] = {}
            for i, p in enumerate(names):
                x = self.samples[:, i].astype(np.float64)
                var = x.var()
                slope = 0.0 if var == 0 else float(np.cov(x, values, bias=True)[0, 1] / var)
                lo, hi = self.param_ranges[p]
                out[metric][p] = abs(slope) * (hi - lo)
        return out

    def plot(self, metric=None, ax=None, **kwargs):
        plt = _pyplot()
        if ax is None:
            _, ax = plt.subplots()
        indices = self.sobol_indices()
        if metric is None and self.metrics_of_interest:
            metric = self.metrics_of_interest[0]
        if metric in indices:
            items = sorted(indices[metric].items(), key=lambda kv: kv[1], reverse=True)
            ax.bar([k for k, _ in items], [v for _, v in items], **kwargs)
            ax.set_xlabel("Parameter")
            ax.set_ylabel("Sensitivity Index")
            ax.set_title(f"Sensitivity Indices for {metric}")
        return ax


class ModelCalibrator:
    RL_METHODS = ("q_learning", "policy_gradient", "actor_critic", "multi_agent_rl", "dqn")

    def __init__(self, model_factory: ModelFactory, initial_params: Dict[str, float],
                 target_metrics: Dict[str, float],
                 param_bounds: Optional[Dict[str, Tuple[float, float]]] = None,
                 metrics_weights: Optional[Dict[str, float]] = None, learning_rate: float = 0.01,
                 max_iterations: int = 100, method: str = "adam", loss_type: str = "mse",
                 evaluation_steps: int = 50, num_evaluation_runs: int = 3, tolerance: float = 1e-6,
                 patience: int = 10, seed: int = 0, fd_step: float = 1e-3):
        if method in self.RL_METHODS:
            raise NotImplementedError(f"calibration method '{method}' (reinforcement learning) is "
                                      "not part of this engine; use adam, sgd, es, pso or cem")
        if method not in ("adam", "sgd", "es", "pso", "cem"):
            raise ValueError(f"Unknown calibration method: {method}")
        self.model_factory = model_factory
        self.params = dict(initial_params)
        self.target_metrics = target_metrics
        self.param_bounds = param_bounds or {k: (0.01, 10.0) for k in initial_params}
        self.metrics_weights = metrics_weights or {k: 1.0 for k in target_metrics}
        self.learning_rate = learning_rate
        self.max_iterations = max_iterations
        self.method = method
        self.loss_type = loss_type
        self.evaluation_steps = evaluation_steps
        self.num_evaluation_runs = num_evaluation_runs
        self.tolerance = tolerance
        self.patience = patience
        self.fd_step = fd_step
        self.key = jr.host_prng_key(seed)
        self.loss_history: List[float] = []
        self.param_history: List[Dict[str, float]] = []
        self.confidence_intervals: List[Any] = []
        self.best_params = dict(initial_params)
        self.best_loss = float("inf")
        self._names = list(self.params.keys())
        self.last_gradient: Optional[str] = None

    # -- losses (jaxabm/analysis.py:401-487)
    def _compute_loss(self, metrics: Dict[str, float]) -> float:
        loss = 0.0
        for m, target in self.target_metrics.items():
            if m not in metrics:
                continue
            v, w = metrics[m], self.metrics_weights.get(m, 1.0)
            if self.loss_type == "mse":
                ml = (v - target) ** 2
            elif self.loss_type == "mae":
                ml = abs(v - target)
            elif self.loss_type == "huber":
                r = abs(v - target)
                ml = 0.5 * r * r if r <= 1.0 else (r - 0.5)
            elif self.loss_type == "relative":
                ml = abs(v - target) / (abs(target) + 1e-8)
            else:
                raise ValueError(f"Unknown loss type: {self.loss_type}")
            loss += w * ml
        return float(loss)

    def _compute_normalized_loss(self, metrics: Dict[str, float]) -> float:
        total = 0.0
        for m, target in self.target_metrics.items():
            if m in metrics:
                total += (abs(metrics[m] - target) / (abs(target) + 1e-8)) ** 2
        return float(total)

    def _metrics_many(self, runs: List[Tuple[Dict[str, float], int]]) -> List[Dict[str, float]]:
        """Target metrics of each (params, seed) run after evaluation_steps
        steps.  Library models with a server request run as one continuous
        batch (same results as one after another)."""
        models = [self.model_factory(params=p, config=ModelConfig(seed=sd)) for p, sd in runs]
        if len(models) > 1 and all(_servable(m) for m in models):
            results = _run_models(models, self.evaluation_steps)
        else:
            results = [m.run(steps=self.evaluation_steps) for m in models]
        out = []
        for res in results:
            d = {}
            for m in self.target_metrics:
                v = _last(res, m)
                d[m] = 0.0 if v is None else v
            out.append(d)
        return out

    def _metrics(self, params: Dict[str, float], seed: int) -> Dict[str, float]:
        return self._metrics_many([(params, seed)])[0]

    def _robust_seeds(self) -> List[int]:
        seeds = []
        for _ in range(self.num_evaluation_runs):
            self.key, sub = jr.host_split(self.key, 2)
            seeds.append(int(jr.randint(sub, (), 0, 1_000_000).item()))
        return seeds

    def _robust_summary(self, metric_runs: List[Dict[str, float]]):
        runs = {m: [r[m] for r in metric_runs] for m in self.target_metrics}
        mean, ci = {}, {}
        for m, vals in runs.items():
            a = np.asarray(vals, np.float64)
            mean[m] = float(a.mean())
            half = 1.96 * float(a.std()) / np.sqrt(len(a))
            ci[m] = (mean[m] - half, mean[m] + half)
        return self._compute_normalized_loss(mean), ci

    def _evaluate_params_robust(self, params):
        """Mean metrics over num_evaluation_runs seeds drawn as
        randint(sub, (), 0, 1_000_000); normalized loss and 95% intervals."""
        seeds = self._robust_seeds()
        return self._robust_summary(self._metrics_many([(params, sd) for sd in seeds]))

    def _clip(self, params: Dict[str, float]) -> Dict[str, float]:
        return {k: float(np.clip(v, *self.param_bounds.get(k, (-np.inf, np.inf))))
                for k, v in params.items()}

    def loss_fn(self, params: Dict[str, float]) -> float:
        """Loss of one run at the fixed gradient seed 42."""
        return self._compute_loss(self._metrics(params, 42))

    def _loss_tensor(self, metrics: Dict[str, torch.Tensor]):
        """_compute_loss on 0-d tensors, differentiably."""
        loss = None
        for m, target in self.target_metrics.items():
            if m not in metrics:
                continue
            d, w = metrics[m] - target, self.metrics_weights.get(m, 1.0)
            if self.loss_type == "mse":
                ml = d * d
            elif self.loss_type == "mae":
                ml = d.abs()
            elif self.loss_type == "huber":
                r = d.abs()
                ml = torch.where(r <= 1.0, 0.5 * r * r, r - 0.5)
            elif self.loss_type == "relative":
                ml = d.abs() / (abs(target) + 1e-8)
            else:
                raise ValueError(f"Unknown loss type: {self.loss_type}")
            loss = w * ml if loss is None else loss + w * ml
        return loss

    def _autograd_grad(self, params: Dict[str, float]) -> Optional[np.ndarray]:
        """d loss / d params by torch autograd through one eager run at seed
        42, the counterpart of the reference's jax.grad of the unrolled run
        (jaxabm/analysis.py:489-531).  The parameters enter the model as
        float32 tensors.  None -- and the caller differentiates numerically
        -- when the run does not keep them as tensors all the way to the
        target metrics: a host read of a parameter (float(), .item(), a
        branch on it; JAX would refuse it too), a metric that is not a
        tensor, or a step that fails on tensor parameters."""
        dev = jr._device()
        pt = {k: torch.tensor(float(v), dtype=torch.float32, device=dev, requires_grad=True)
              for k, v in params.items()}
        try:
            with torch.enable_grad(), _NoConcretization():
                model = self.model_factory(params=pt, config=ModelConfig(seed=42))
                if _servable(model):
                    return None
                res = model.run(steps=self.evaluation_steps)
                data = getattr(res, "_raw", res)
                metrics = {}
                for m in self.target_metrics:
                    v = data.get(m) if isinstance(data, dict) else None
                    if v is None:
                        continue
                    last = v[-1] if isinstance(v, (list, tuple)) else v
                    if not isinstance(last, torch.Tensor) or last.numel() != 1:
                        return None
                    metrics[m] = last.float().reshape(())
                loss = self._loss_tensor(metrics)
                if loss is None or not loss.requires_grad:
                    return None
                grads = torch.autograd.grad(loss, [pt[k] for k in self._names], allow_unused=True)
        except (_Concrete, TypeError, RuntimeError, ValueError, AttributeError, KeyError):
            return None
        return np.array([0.0 if g is None else float(g) for g in grads], np.float64)

    def grad_fn(self, params: Dict[str, float]) -> np.ndarray:
        """d loss / d params at seed 42: the model's own adjoint when it has
        ``loss_gradient``, else torch autograd through the run, else central
        differences (``last_gradient`` says which)."""
        model = self.model_factory(params=params, config=ModelConfig(seed=42))
        if hasattr(model, "loss_gradient"):
            self.last_gradient = "adjoint"
            g = model.loss_gradient(self.target_metrics, self.metrics_weights,
                                    steps=self.evaluation_steps)
            return np.array([g[k] for k in self._names], np.float64)
        if not _servable(model):
            g = self._autograd_grad(params)
            if g is not None:
                self.last_gradient = "autograd"
                return g
        self.last_gradient = "finite_difference"
        g = np.zeros(len(self._names))
        hs, runs = [], []
        for k in self._names:
            h = self.fd_step * max(1.0, abs(params[k]))
            hs.append(h)
            runs += [({**params, k: params[k] + h}, 42), ({**params, k: params[k] - h}, 42)]
        losses = [self._compute_loss(m) for m in self._metrics_many(runs)]
        for i, h in enumerate(hs):
            g[i] = (losses[2 * i] - losses[2 * i + 1]) / (2 * h)
        return g

    # -- optimisers
    def _record(self, params, loss):
        self.loss_history.append(float(loss))
        self.param_history.append(dict(params))
        if loss < self.best_loss:
            self.best_loss, self.best_params = float(loss), dict(params)

    def _calibrate_gradient(self, verbose: bool) -> Dict[str, float]:
        p = np.array([self.params[k] for k in self._names], np.float64)
        m = np.zeros_like(p)
        v = np.zeros_like(p)
        b1, b2, eps = 0.9, 0.999, 1e-8
        stall = 0
        for it in range(1, self.max_iterations + 1):
            params = dict(zip(self._names, p.tolist()))
            loss = self.loss_fn(params)
            prev = self.best_loss
            self._record(params, loss)
            stall = stall + 1 if prev - self.best_loss < self.tolerance else 0
            if verbose:
                print(f"iteration {it}: loss {loss:.6g} params {params}")
            if stall >= self.patience:
                break
            g = self.grad_fn(params)
            if self.method == "adam":
                m = b1 * m + (1 - b1) * g
                v = b2 * v + (1 - b2) * g * g
                step = self.learning_rate * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + eps)
            else:
                step = self.learning_rate * g
            p = np.array(list(self._clip(dict(zip(self._names, (p - step).tolist()))).values()))
        self.params = self.best_params
        return dict(self.best_params)

    def _normal(self, shape) -> np.ndarray:
        self.key, sub = jr.host_split(self.key, 2)
        return jr.normal(sub, shape).double().cpu().numpy()

    def _uniform(self, shape) -> np.ndarray:
        self.key, sub = jr.host_split(self.key, 2)
        return jr.uniform(sub, shape).double().cpu().numpy()

    def _population_loss(self, pop: np.ndarray) -> np.ndarray:
        """Robust loss of every member; the seeds are drawn member by member
        as the serial loop draws them, and all member x seed runs go out as
        one batch."""
        members = [self._clip(dict(zip(self._names, x))) for x in pop]
        seeds = [self._robust_seeds() for _ in members]
        flat = self._metrics_many([(p, sd) for p, ss in zip(members, seeds) for sd in ss])
        k = self.num_evaluation_runs
        return np.array([self._robust_summary(flat[i * k:(i + 1) * k])[0]
                         for i in range(len(members))])

    def _bounds_arrays(self):
        lo = np.array([self.param_bounds.get(k, (-np.inf, np.inf))[0] for k in self._names])
        hi = np.array([self.param_bounds.get(k, (-np.inf, np.inf))[1] for k in self._names])
        return lo, hi

    def _calibrate_es(self, verbose: bool, pop_size=20, sigma=0.1, elite=0.2):
        lo, hi = self._bounds_arrays()
        mean = np.array([self.params[k] for k in self._names], np.float64)
        n_elite = max(1, int(pop_size * elite))
        for it in range(self.max_iterations):
            pop = np.clip(mean + sigma * self._normal((pop_size, len(mean))), lo, hi)
            losses = self._population_loss(pop)
            order = np.argsort(losses, kind="stable")
            mean = pop[order[:n_elite]].mean(0)
            self._record(dict(zip(self._names, pop[order[0]].tolist())), losses[order[0]])
            if verbose:
                print(f"iteration {it + 1}: best loss {losses[order[0]]:.6g}")
        return dict(self.best_params)

    def _calibrate_pso(self, verbose: bool, pop_size=20, w=0.7, c1=1.5, c2=1.5):
        lo, hi = self._bounds_arrays()
        d = len(self._names)
        span = np.where(np.isfinite(hi - lo), hi - lo, 1.0)
        x = np.clip(np.array([self.params[k] for k in self._names]) +
                    0.1 * span * self._normal((pop_size, d)), lo, hi)
        vel = np.zeros_like(x)
        pbest, pbest_loss = x.copy(), self._population_loss(x)
        g = pbest[np.argmin(pbest_loss)].copy()
        for it in range(self.max_iterations):
            r1, r2 = self._uniform((pop_size, d)), self._uniform((pop_size, d))
            vel = w * vel + c1 * r1 * (pbest - x) + c2 * r2 * (g - x)
            x = np.clip(x + vel, lo, hi)
            losses = self._population_loss(x)
            better = losses < pbest_loss
            pbest[better], pbest_loss[better] = x[better], losses[better]
            g = pbest[np.argmin(pbest_loss)].copy()
            self._record(dict(zip(self._names, g.tolist())), float(pbest_loss.min()))
            if verbose:
                print(f"iteration {it + 1}: best loss {pbest_loss.min():.6g}")
        return dict(self.best_params)

    def _calibrate_cem(self, verbose: bool, pop_size=20, elite=0.2):
        lo, hi = self._bounds_arrays()
        mean = np.array([self.params[k] for k in self._names], np.float64)
        span = np.where(np.isfinite(hi - lo), hi - lo, 1.0)
        std = 0.25 * span
        n_elite = max(1, int(pop_size * elite))
        for it in range(self.max_iterations):
            pop = np.clip(mean + std * self._normal((pop_size, len(mean))), lo, hi)
            losses = self._population_loss(pop)
            order = np.argsort(losses, kind="stable")
            el = pop[order[:n_elite]]
            mean, std = el.mean(0), el.std(0) + 1e-6
            self._record(dict(zip(self._names, pop[order[0]].tolist())), losses[order[0]])
            if verbose:
                print(f"iteration {it + 1}: best loss {losses[order[0]]:.6g}")
        return dict(self.best_params)

    def calibrate(self, verbose: bool = True) -> Dict[str, float]:
        if self.method in ("adam", "sgd"):
            return self._calibrate_gradient(verbose)
        if self.method == "es":
            return self._calibrate_es(verbose)
        if self.method == "pso":
            return self._calibrate_pso(verbose)
        return self._calibrate_cem(verbose)

    def get_calibration_history(self) -> Dict[str, List[Any]]:
        return {"loss": list(self.loss_history), "params": list(self.param_history),
                "confidence_intervals": list(self.confidence_intervals)}
