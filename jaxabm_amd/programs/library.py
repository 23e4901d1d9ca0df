"""The library programs as façade models: ``model.run()`` returns
``Results`` like any ``jaxabm_amd.agentpy.Model``, but the whole run is one
fused HIP engine instead of per-agent vmapped callbacks.

Each accepts ``(parameters, seed=None, config=None)`` so it can be built by
the reference's model factories (``factory(params=..., config=...)``,
jaxabm/analysis.py:123-127) and used by ``SensitivityAnalysis`` /
``ModelCalibrator`` unchanged.  Results hold ``'step'`` (1..steps) and the
program's records as per-step lists.
"""
from __future__ import annotations

import time
from typing import Any, Dict, Optional

import numpy as np

from ..agentpy import Model, Results
from ..core import ModelConfig
from ..utils import format_time


class _ProgramModel(Model):
    defaults: Dict[str, Any] = {}

    def __init__(self, parameters: Optional[Dict[str, Any]] = None, seed: Optional[int] = None,
                 config: Optional[ModelConfig] = None):
        p = {**self.defaults, **(parameters or {})}
        if seed is None and config is not None:
            seed = config.seed
        super().__init__(p, seed)
        self.engine = None

    def _results(self, steps: int, records: Dict[str, Any], t0: float) -> Results:
        data = {"step": list(range(1, steps + 1))}
        data.update({k: np.asarray(v).tolist() for k, v in records.items()})
        data.update(self._recorded_data)
        print(f"Simulation executed in {format_time(time.time() - t0)}")
        return Results(data)


class WealthModel(_ProgramModel):
    """Wealth-exchange random walk: parameters ``n`` (1000), ``steps`` (100);
    records ``mean_wealth`` and ``gini``."""

    defaults = {"n": 1000, "steps": 100}

    def serve_request(self, steps: Optional[int] = None):
        """The same run as a request to the batching server."""
        from ..serving import SimRequest

        steps = int(steps if steps is not None else self.p.get("steps", 100))
        return SimRequest("wealth", n=int(self.p["n"]), steps=steps, seed=int(self.seed))

    def results_from_records(self, steps: int, rec: Dict[str, Any], t0: float) -> Results:
        return self._results(steps, {"mean_wealth": rec["mean"], "gini": rec["gini"]}, t0)

    def run(self, steps: Optional[int] = None) -> Results:
        from .wealth import WealthEngine

        steps = int(steps if steps is not None else self.p.get("steps", 100))
        t0 = time.time()
        self.engine = WealthEngine(int(self.p["n"]), seed=int(self.seed))
        mean, gini = self.engine.run(steps)
        return self._results(steps, {"mean_wealth": mean, "gini": gini}, t0)


class SchellingModel(_ProgramModel):
    """Schelling segregation on a periodic ``height`` x ``width`` Grid (512 x
    512), ``steps`` (200); records ``fraction_happy`` and ``moves``."""

    defaults = {"height": 512, "width": 512, "steps": 200}

    def run(self, steps: Optional[int] = None) -> Results:
        from .schelling import SchellingEngine

        steps = int(steps if steps is not None else self.p.get("steps", 200))
        t0 = time.time()
        self.engine = SchellingEngine(int(self.p["height"]), int(self.p["width"]),
                                      seed=int(self.seed))
        happy, moves = self.engine.run(steps)
        return self._results(steps, {"fraction_happy": happy, "moves": moves}, t0)


class NetworkSIRModel(_ProgramModel):
    """SIR on a Barabasi-Albert network: ``n`` (1,000,000), ``m`` (4),
    ``beta`` (0.05), ``gamma`` (0.1), ``n_infected`` (100), ``graph_seed`` (0),
    ``steps`` (365); records ``S``, ``I``, ``R``."""

    defaults = {"n": 1_000_000, "m": 4, "beta": 0.05, "gamma": 0.1, "n_infected": 100,
                "graph_seed": 0, "steps": 365}

    def serve_request(self, steps: Optional[int] = None):
        """The same run as a request to the batching server."""
        from ..serving import SimRequest

        steps = int(steps if steps is not None else self.p.get("steps", 365))
        return SimRequest("sir", n=int(self.p["n"]), steps=steps, seed=int(self.seed),
                          graph_seed=int(self.p["graph_seed"]), m=int(self.p["m"]),
                          n_infected=int(self.p["n_infected"]), beta=float(self.p["beta"]),
                          gamma=float(self.p["gamma"]))

    def results_from_records(self, steps: int, rec: Dict[str, Any], t0: float) -> Results:
        return self._results(steps, {"S": rec["S"], "I": rec["I"], "R": rec["R"]}, t0)

    def run(self, steps: Optional[int] = None) -> Results:
        from .sir import SIRNetworkEngine

        steps = int(steps if steps is not None else self.p.get("steps", 365))
        t0 = time.time()
        # the BA graph is built on the device by the engine
        self.engine = SIRNetworkEngine(int(self.p["n"]), m=int(self.p["m"]),
                                       graph_seed=int(self.p["graph_seed"]),
                                       beta=float(self.p["beta"]),
                                       gamma=float(self.p["gamma"]),
                                       n_seed_infected=int(self.p["n_infected"]),
                                       seed=int(self.seed))
        c = self.engine.run(steps)
        return self._results(steps, {"S": c[:, 0], "I": c[:, 1], "R": c[:, 2]}, t0)


DEFFUANT_METRICS = ["opinion_mean", "opinion_var"] + [f"hist_{k}" for k in range(32)]


class DeffuantOpinionModel(_ProgramModel):
    """Differentiable Deffuant model: ``n`` (65,536), ``eps`` (0.2), ``mu``
    (0.3), ``steps`` (200); records the final ``opinion_mean``,
    ``opinion_var`` and ``hist_0`` .. ``hist_31`` (one value each).
    :meth:`loss_gradient` gives d(loss)/d(eps, mu) of a weighted squared error
    against target metrics from the hand-written adjoint."""

    defaults = {"n": 65536, "eps": 0.2, "mu": 0.3, "steps": 200, "ckpt_every": 20}

    def _model(self, steps):
        from .deffuant import DeffuantModel

        return DeffuantModel(int(self.p["n"]), 1, steps, seed=int(self.seed),
                             ckpt_every=int(self.p["ckpt_every"]))

    def run(self, steps: Optional[int] = None) -> Results:
        steps = int(steps if steps is not None else self.p.get("steps", 200))
        t0 = time.time()
        self.engine = self._model(steps)
        self.engine.forward([self.p["eps"]], [self.p["mu"]], np.zeros(34))
        st = self.engine.stats[0].cpu().numpy()
        print(f"Simulation executed in {format_time(time.time() - t0)}")
        data = {"step": [steps]}
        data.update({name: [float(st[i])] for i, name in enumerate(DEFFUANT_METRICS)})
        data.update(self._recorded_data)
        return Results(data)

    def loss_gradient(self, target_metrics: Dict[str, float], metrics_weights: Dict[str, float],
                      steps: Optional[int] = None) -> Dict[str, float]:
        steps = int(steps if steps is not None else self.p.get("steps", 200))
        target = np.zeros(34)
        weights = np.zeros(34)
        for i, name in enumerate(DEFFUANT_METRICS):
            if name in target_metrics:
                target[i] = target_metrics[name]
                weights[i] = metrics_weights.get(name, 1.0)
        eng = self._model(steps)
        eng.forward([self.p["eps"]], [self.p["mu"]], target, weights)
        ge, gm = eng.backward()
        return {"eps": float(ge[0]), "mu": float(gm[0])}


def servable(model) -> bool:
    """Whether ``model`` runs as a request of the batching server."""
    return callable(getattr(model, "serve_request", None)) and \
        callable(getattr(model, "results_from_records", None))


def run_models(models, steps: Optional[int] = None, config: Optional[Dict[str, Any]] = None):
    """``[m.run(steps) for m in models]`` with the same results, where the
    models that have a server request (wealth, network SIR) are submitted
    together to one continuous-batching server instead of run one after
    another (the serial loops of jaxabm/analysis.py:97-165, 434-476,
    jaxabm/utils.py:128-176 and jaxabm/agentpy.py:1116-1152).  A request the
    server fails (e.g. larger than its memory budget) runs on its own."""
    from .. import random as jr
    from ..serving import FAILED, SimulationServer, tuned_config

    out: list = [None] * len(models)
    reqs = {i: m.serve_request(steps) for i, m in enumerate(models) if servable(m)}
    if len(reqs) >= 2:
        cfg = tuned_config("batched")
        cfg["mode"] = jr.default_mode()
        cfg.update(config or {})
        t0 = time.time()
        with SimulationServer(cfg) as server:
            tickets = {i: server.submit(r) for i, r in reqs.items()}
            server.drain()
            for i, t in tickets.items():
                if server.info(t).status != FAILED:
                    out[i] = models[i].results_from_records(reqs[i].steps, server.result(t), t0)
        print(f"{sum(o is not None for o in out)} simulations executed together in "
              f"{format_time(time.time() - t0)}")
    for i, m in enumerate(models):
        if out[i] is None:
            out[i] = m.run(steps)
    return out
