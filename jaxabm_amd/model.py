"""The model driver (jaxabm/model.py:17-341): agent collections, environment
state, the step schedule and history.

Step t (jaxabm/model.py:146-216):
  rng, step_key = split(rng)
  model_state = {time_step, env, agents_<name>...} (before any update)
  for each collection in insertion order: step_key, ck = split(step_key);
      collection.update(model_state, ck, config)
  if update_state_fn: step_key, uk = split(step_key);
      env = update_state_fn(env, agent_states, params, uk)
  metrics = metrics_fn(env, agent_states, params); time_step += 1; history.
The key arithmetic runs on the host (two words per split, no device sync);
keys handed to user code are device tensors of shape (2,).
"""
from __future__ import annotations

import time
from typing import Any, Callable, Dict, List, Optional

import numpy as np
import torch

from . import numpy as jnp
from . import random as jr
from . import reductions
from .agent import AgentCollection, _fn_id, vmap_agents
from .core import ModelConfig


def _dev_key(k: np.ndarray) -> torch.Tensor:
    """A host key on the device, staged in pinned memory and copied on the
    current stream without waiting for the work queued before it."""
    host = torch.from_numpy(np.ascontiguousarray(k, np.uint32).view(np.int32).copy())
    return host.pin_memory().to(jr._device(), non_blocking=True)


class Model:
    """Core model (the reference's ``JaxModel``)."""

    def __init__(self, params: Optional[Dict[str, Any]] = None, config: Optional[ModelConfig] = None,
                 update_state_fn: Optional[Callable] = None, metrics_fn: Optional[Callable] = None):
        self.config = config or ModelConfig()
        self._rng = jr.host_prng_key(self.config.seed)
        self._agent_collections: Dict[str, AgentCollection] = {}
        self._env_state: Dict[str, Any] = {}
        self._state: Optional[Dict[str, Any]] = None
        self._params = params or {}
        # the callbacks see JAX's promotion of booleans in subtraction and
        # JAX's gather indexing
        self._update_state_fn = jnp.with_jax_semantics(update_state_fn)
        self._metrics_fn = jnp.with_jax_semantics(metrics_fn)
        self._time_step = 0
        self._history: List[Dict[str, Any]] = []
        self._is_initialized = False
        self.last_run_info: Dict[str, Any] = {}

    def add_agent_collection(self, name: str, agent_collection: AgentCollection) -> None:
        if self._is_initialized:
            raise RuntimeError("Cannot add agent collections after model is initialized")
        self._agent_collections[name] = agent_collection

    def add_env_state(self, name: str, value: Any) -> None:
        """Before initialize: the env dict.  After: both the env dict and the
        state snapshot taken at initialize (jaxabm/model.py:76-99)."""
        if self._is_initialized and self._state is not None:
            self._state.setdefault("env", {})[name] = value
        if self._env_state is None:
            self._env_state = {}
        self._env_state[name] = value

    def model_state(self) -> Dict[str, Any]:
        state = {"time_step": self._time_step, "env": self._env_state}
        for name, coll in self._agent_collections.items():
            state[f"agents_{name}"] = coll.states
        return state

    def initialize(self) -> None:
        if not self._agent_collections:
            raise ValueError("No agent collections added to model")
        keys = jr.host_split(self._rng, len(self._agent_collections) + 1)
        self._rng = keys[0].copy()
        for i, coll in enumerate(self._agent_collections.values()):
            if getattr(coll, "model_config", None) is None:
                coll.model_config = self.config
            coll.init(_dev_key(keys[i + 1]), self.config)
        self._is_initialized = True
        self._state = {"env": self._env_state.copy()}

    def step(self) -> Dict[str, Any]:
        if not self._is_initialized:
            raise RuntimeError("Model must be initialized before stepping. Call initialize() first.")
        self._rng, step_key = jr.host_split(self._rng, 2)
        current = self.model_state()
        for coll in self._agent_collections.values():
            step_key, coll_key = jr.host_split(step_key, 2)
            coll.update(current, _dev_key(coll_key), self.config)
        updated = {name: c.states for name, c in self._agent_collections.items()}
        metrics = {}
        # the callbacks' whole-array sums and means go out as one launch
        with reductions.deferring():
            if self._update_state_fn:
                step_key, update_key = jr.host_split(step_key, 2)
                self._env_state = self._update_state_fn(self._env_state, updated, self._params,
                                                        _dev_key(update_key))
            if self._metrics_fn:
                metrics = self._metrics_fn(self._env_state, updated, self._params)
        self._env_state = reductions.materialize(self._env_state)
        metrics = reductions.materialize(metrics)
        self._time_step += 1
        if self.config.track_history and self._time_step % self.config.collect_interval == 0:
            self._history.append({"time_step": self._time_step, "metrics": metrics})
        return metrics

    def run(self, steps: Optional[int] = None, graph: bool = False) -> Dict[str, List[Any]]:
        """Initialise if needed, step, and return the history as columns
        ({'step': [...], metric: [...]}); a second run continues the state and
        resets only the history (jaxabm/model.py:218-262).

        ``graph=True`` runs steps 2..n as replays of one captured hipGraph of
        the step (same results; ``last_run_info`` says whether it was
        replayed and why not); a step that reads device values on the host
        runs eagerly as usual."""
        if not self._is_initialized:
            self.initialize()
        n = steps if steps is not None else self.config.steps
        if self.config.track_history:
            self._history = []
        t0 = time.time()
        done = 0
        self.last_run_info = {"mode": "eager", "steps": n, "steps_done": 0}
        if graph and n > 1:
            from .graphrun import run_graph

            self.last_run_info = run_graph(self, n)
            done = self.last_run_info["steps_done"]
        for _ in range(n - done):
            self.step()
        elapsed = time.time() - t0
        print(f"Ran {n} steps in {elapsed:.2f}s ({n / max(elapsed, 1e-12):.1f} steps/sec)")
        if self.config.track_history and self._history:
            out = {"step": [h["time_step"] for h in self._history]}
            for name in self._history[0]["metrics"].keys():
                out[name] = [h["metrics"][name] for h in self._history]
            return out
        return {}

    @property
    def agent_collections(self) -> Dict[str, AgentCollection]:
        return self._agent_collections

    @property
    def state(self) -> Dict[str, Any]:
        if self._state is None:
            self._state = {"env": self._env_state.copy()}
        return self._state

    def jit_step(self) -> Callable:
        """A pure step function ``(env, agent_states, params, key) -> (env,
        agent_states, metrics, key)`` with the same schedule as ``step``
        (jaxabm/model.py:264-341).  Nothing is traced: the HIP kernels run as
        they are called."""

        def _step_fn(env_state, agent_states, params, key):
            model_state = {"time_step": self._time_step, "env": env_state}
            updated = {}
            for name, states in agent_states.items():
                coll = self._agent_collections[name]
                key, sub = jr.split(key)
                n = next(iter(states.values())).shape[0]
                upd = coll.agent_type.update
                updated[name] = vmap_agents(
                    lambda s, k: upd(s, model_state, self.config, k), states, jr.split(sub, n), n=n,
                    ident=_fn_id(upd))
            with reductions.deferring():
                if self._update_state_fn:
                    key, uk = jr.split(key)
                    env_state = self._update_state_fn(env_state, updated, params, uk)
                metrics = self._metrics_fn(env_state, updated, params) if self._metrics_fn else {}
            return (reductions.materialize(env_state), updated, reductions.materialize(metrics),
                    key)

        return _step_fn


JaxModel = Model
