"""Agent types and collections (jaxabm/agent.py:19-244).

A collection holds its agents' state as a dict of GPU tensors of shape
(num_agents, ...) (structure of arrays).  ``init`` and ``update`` map the
agent type's per-agent functions over the agents, as the reference does with
``jax.vmap`` (jaxabm/agent.py:115-177), with the same key schedule:
``split(key, num_agents)`` gives agent i its key, and the draws of
``jaxabm_amd.random`` reproduce JAX's vmapped bits.  A function of per-agent
scalars, env scalars and draws runs as one generated kernel
(``jaxabm_amd.fuse``); anything else runs under ``torch.func.vmap``, with
the same results.
"""
from __future__ import annotations

from typing import Any, Callable, Dict, Optional, Protocol

import numpy as np
import torch

from . import numpy as jnp
from . import random as jr
from .core import ModelConfig

AgentState = Dict[str, Any]


class AgentType(Protocol):
    """Protocol of an agent type (jaxabm/agent.py:19-51)."""

    def init_state(self, model_config: Any, key: Any) -> AgentState:
        ...

    def update(self, state: AgentState, model_state: Dict[str, Any], model_config: Any,
               key: Any) -> AgentState:
        ...


def as_state_value(v) -> torch.Tensor:
    """A per-agent output as a tensor: Python scalars take JAX's default
    dtypes (bool, int32, float32), float64 arrays become float32."""
    if isinstance(v, torch.Tensor):
        return v
    # scalars by a fill kernel, not a host copy: no host synchronisation, so a
    # step that returns constants can be captured as a graph
    if isinstance(v, bool):
        return torch.full((), v, dtype=torch.bool, device=jr._device())
    if isinstance(v, int):
        return torch.full((), v, dtype=torch.int32, device=jr._device())
    if isinstance(v, float):
        return torch.full((), v, dtype=torch.float32, device=jr._device())
    a = np.asarray(v)
    if a.dtype == np.float64:
        a = a.astype(np.float32)
    elif a.dtype == np.int64:
        a = a.astype(np.int32)
    return torch.as_tensor(a, device=jr._device())


def _tensor_dict(out, what: str) -> Dict[str, torch.Tensor]:
    if not isinstance(out, dict):
        raise ValueError(f"{what} must return a dictionary, got {type(out)}")
    return {k: as_state_value(v) for k, v in out.items()}


_JAX_SEMANTICS_FNS: set = set()   # agent functions that need jnp.JaxSemantics


def _fn_id(fn):
    f = getattr(fn, "__func__", fn)
    return getattr(f, "__code__", None) or id(f)


def vmap_agents(fn: Callable, *args, n: int, ident: Any = None) -> Dict[str, torch.Tensor]:
    """``vmap(fn)(*args)`` over the leading (agent) axis; outputs that do not
    depend on the agent are broadcast to n agents.

    A function that torch refuses where JAX has a meaning -- a boolean
    operand of a subtraction, or an agents array indexed by a per-agent
    integer -- is run again under ``jnp.JaxSemantics`` and remembered by
    ``ident`` (agent functions are pure, as under ``jax.vmap``, so the second
    run is the first's)."""
    mapped = torch.func.vmap(lambda *a: _tensor_dict(fn(*a), "agent function"))
    key = ident if ident is not None else _fn_id(fn)
    if key in _JAX_SEMANTICS_FNS:
        with jnp.JaxSemantics():
            out = mapped(*args)
    else:
        try:
            out = mapped(*args)
        except RuntimeError as e:
            if not any(m in str(e) for m in jnp.JAX_SEMANTICS_ERRORS):
                raise
            _JAX_SEMANTICS_FNS.add(key)
            with jnp.JaxSemantics():
                out = mapped(*args)
    return {k: (v if v.dim() and v.shape[0] == n else v.expand(n, *v.shape)).contiguous()
            for k, v in out.items()}


class AgentCollection:
    """Collection of agents of one type (jaxabm/agent.py:54-244)."""

    def __init__(self, agent_type: AgentType, num_agents: int):
        if not isinstance(num_agents, int) or isinstance(num_agents, bool) or num_agents <= 0:
            raise ValueError("num_agents must be a positive integer")
        self.agent_type = agent_type
        self.num_agents = num_agents
        self.model_config: Optional[ModelConfig] = None
        self._key = None
        self._states: Optional[Dict[str, torch.Tensor]] = None

    def _name(self) -> str:
        t = self.agent_type
        return getattr(t, "__name__", type(t).__name__)

    def init(self, key: Any, model_config: ModelConfig) -> None:
        """Agent keys = split(key, num_agents); states = vmap(init_state)."""
        if not isinstance(model_config, ModelConfig):
            raise TypeError("model_config must be a ModelConfig instance.")
        if not isinstance(self.num_agents, int) or self.num_agents <= 0:
            raise ValueError("Number of agents must be a positive integer.")
        self._key = key
        self.model_config = model_config
        init_method = getattr(self.agent_type, "init_state", None)
        if not callable(init_method):
            raise AttributeError(f"Agent type {self._name()} must implement 'init_state'")
        from . import fuse

        out = fuse.run_agents(init_method, self.num_agents, key, jr.default_mode(), None, None,
                              lambda s, ms, k: init_method(model_config, k))
        if out is not None:
            self._states = out
            return
        agent_keys = jr.split(key, self.num_agents)
        self._states = vmap_agents(lambda k: init_method(model_config, k), agent_keys,
                                   n=self.num_agents, ident=_fn_id(init_method))

    def update(self, model_state: Dict[str, Any], key: Any, model_config: ModelConfig) -> None:
        """Agent keys = split(key, num_agents); states = vmap(update)."""
        if self._states is None:
            raise ValueError("Agent collection not initialized. Call init() first.")
        if self.model_config is None:
            raise RuntimeError("Model config not set for AgentCollection. Ensure Model.initialize() "
                               "was called.")
        update_method = getattr(self.agent_type, "update", None)
        if not callable(update_method):
            raise AttributeError(f"Agent type {self._name()} must implement 'update'")
        from . import fuse

        out = fuse.run_agents(update_method, self.num_agents, key, jr.default_mode(),
                              self._states, model_state,
                              lambda s, ms, k: update_method(s, ms, model_config, k))
        if out is not None:
            self._states = out
            return
        agent_keys = jr.split(key, self.num_agents)
        self._states = vmap_agents(
            lambda s, k: update_method(s, model_state, model_config, k), self._states, agent_keys,
            n=self.num_agents, ident=_fn_id(update_method))

    def get_states(self) -> Dict[str, Any]:
        return self._states

    @property
    def states(self) -> Dict[str, torch.Tensor]:
        return self._states

    def aggregate(self, variable: str, fn: Callable = None) -> Any:
        """Reduce one state variable over the agents (default: mean)."""
        if self._states is None or variable not in self._states:
            raise ValueError(f"Variable {variable} not found in agent states")
        if fn is None:
            from . import numpy as jnp

            fn = jnp.mean
        return fn(self._states[variable])

    def filter(self, condition: Callable[[Dict[str, Any]], Any]) -> "AgentCollection":
        """New collection of the agents whose mask is true (same type, config
        and key); its size is read back to the host."""
        mask = condition({k: self._states[k] for k in self._states})
        mask = mask if isinstance(mask, torch.Tensor) else torch.as_tensor(mask)
        mask = mask.to(device=next(iter(self._states.values())).device, dtype=torch.bool)
        count = int(mask.sum().item())
        filtered = AgentCollection(agent_type=self.agent_type, num_agents=count)
        filtered.model_config = self.model_config
        filtered._key = self._key
        filtered._states = {k: v[mask] for k, v in self._states.items()}
        return filtered
