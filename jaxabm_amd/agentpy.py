"""AgentPy-style façade (jaxabm/agentpy.py:1-1485): Agent, AgentList,
Environment, Grid, Network, Model, Results, Parameter, Sample,
SensitivityAnalyzer, ModelCalibrator.

Semantics follow the reference, including its observable quirks:
  * ``Agent.setup()`` takes no key and its output is the same for every agent
    (broadcast); ``Agent.step(model_state)`` takes no key
    (jaxabm/agentpy.py:175-227);
  * the env phase runs ``Model.step()`` eagerly and merges
    ``{**env_state, **self.env.state}``: values set on the Environment
    override the env phase every step (jaxabm/agentpy.py:895-924);
  * ``Grid.position_agents`` is a no-op before the collection is initialised
    (jaxabm/agentpy.py:491-527);
  * undirected ``Network.add_edge`` stores both directions and
    ``get_neighbors`` returns the sorted unique ids (jaxabm/agentpy.py:
    559-615); the edge list is kept on the host and ``network_edges`` is
    materialised when read, so building is O(E) instead of O(E^2), and a bulk
    Barabasi-Albert constructor builds large graphs directly;
  * ``Results`` converts device tensors to numpy and exposes
    ``variables.<type>.<var>`` for ``agents.<type>.<var>`` keys; ``save`` /
    ``load`` pickle the data dict as the reference does.
"""
from __future__ import annotations

import itertools
import time
from typing import Any, Callable, Dict, List, Optional, Tuple, Type, Union

import numpy as np
import torch

from . import random as jr
from .agent import AgentCollection, AgentType
from .core import ModelConfig
from .model import Model as JaxModel
from .utils import convert_to_numpy, format_time


class Agent:
    """Base class of façade agents: override ``setup`` and ``step``."""

    def __init__(self):
        self.id = None
        self.model = None
        self.p = {}
        self._state = {}

    def setup(self) -> Dict[str, Any]:
        return {}

    def step(self, model_state: Optional[Dict[str, Any]] = None) -> Dict[str, Any]:
        return self._state

    def update_state(self, new_state: Dict[str, Any]) -> None:
        if self.model is not None and hasattr(self.model, "_update_agent_state"):
            self.model._update_agent_state(self, new_state)
        else:
            self._state.update(new_state)

    def __getattr__(self, name: str) -> Any:
        state = self.__dict__.get("_state")
        if state and name in state:
            return state[name]
        raise AttributeError(f"'{self.__class__.__name__}' object has no attribute '{name}'")

    def __setattr__(self, name: str, value: Any) -> None:
        if name in ("id", "model", "p", "_state"):
            super().__setattr__(name, value)
        else:
            if "_state" not in self.__dict__:
                super().__setattr__("_state", {})
            self.update_state({name: value})


class AgentWrapper(AgentType):
    """Adapter from an Agent class to the AgentType protocol
    (jaxabm/agentpy.py:157-227)."""

    def __init__(self, agent_class: Type[Agent], params: Optional[Dict[str, Any]] = None):
        self.agent_class = agent_class
        self.params = params or {}
        self.agent_instance = agent_class()
        if params:
            self.agent_instance.p = params

    @property
    def __name__(self) -> str:
        return self.agent_class.__name__

    def init_state(self, model_config: ModelConfig, key: Any) -> Dict[str, Any]:
        state = self.agent_instance.setup()
        if not isinstance(state, dict):
            if state is None:
                return {}
            raise ValueError(f"Agent.setup() must return a dictionary, got {type(state)}")
        return state

    def update(self, state: Dict[str, Any], model_state: Dict[str, Any],
               model_config: ModelConfig, key: Any) -> Dict[str, Any]:
        self.agent_instance._state = state
        new_state = self.agent_instance.step(model_state)
        if not isinstance(new_state, dict):
            if new_state is None:
                return state
            raise ValueError(f"Agent.step() must return a dictionary, got {type(new_state)}")
        return new_state


AgentTypeWrapper = AgentWrapper


class AgentList:
    """A named group of agents of one class (jaxabm/agentpy.py:230-378)."""

    def __init__(self, model: "Model", n: int, agent_class: Type[Agent], **kwargs):
        self.model = model
        self.n = n
        self.agent_class = agent_class
        self.params = kwargs
        self.agent_type = AgentWrapper(agent_class, kwargs)
        self.collection = AgentCollection(agent_type=self.agent_type, num_agents=n)
        self.name = None

    @property
    def states(self) -> Dict[str, Any]:
        jm = getattr(self.model, "_jax_model", None)
        if jm is not None and isinstance(getattr(jm, "state", None), dict):
            agent_states = jm.state.get("agents", {})
            if self.name and self.name in agent_states:
                return agent_states[self.name]
        st = getattr(self.collection, "states", None)
        return st if st is not None else {}

    def __getattr__(self, name: str) -> Any:
        if name.startswith("__") or name in ("model", "collection", "name", "n"):
            raise AttributeError(name)
        states = self.states
        if name in states:
            return states[name]
        raise AttributeError(f"'AgentList' object has no attribute '{name}'")

    def __len__(self) -> int:
        return self.n

    def select(self, condition: Callable[[Any], Any]) -> "AgentList":
        states = self.states
        mask = condition(states)
        mask_t = mask if isinstance(mask, torch.Tensor) else torch.as_tensor(np.asarray(mask))
        count = int(mask_t.sum().item())
        filtered = AgentList(model=self.model, n=count, agent_class=self.agent_class, **self.params)
        if count and self.collection.states is not None:
            filtered.collection = self.collection.filter(lambda s: mask_t)
        return filtered

    def __iter__(self):
        inst = getattr(self.model, "_agent_instances", None) if self.model is not None else None
        if inst and self.name in inst:
            yield from inst[self.name]
            return
        states = self.states
        if not states:
            return
        host = {k: convert_to_numpy(v) for k, v in states.items()}
        count = len(next(iter(host.values())))
        for i in range(count):
            agent = self.agent_class()
            agent.model = self.model
            agent.id = i
            agent.p = self.params
            agent._state = {k: host[k][i] for k in host}
            yield agent


class Environment:
    """Environment state container (jaxabm/agentpy.py:381-462)."""

    def __init__(self, model: "Model"):
        self.model = model
        self.state = {}

    def add_state(self, name: str, value: Any) -> None:
        self.state[name] = value
        jm = getattr(self.model, "_jax_model", None)
        if jm is not None:
            jm.add_env_state(name, value)

    def __getattr__(self, name: str) -> Any:
        if name in ("model", "state"):
            raise AttributeError(name)
        st = self.__dict__.get("state", {})
        if name in st:
            v = st[name]
            return v() if isinstance(v, _Lazy) else v
        jm = getattr(self.__dict__.get("model"), "_jax_model", None)
        if jm is not None and jm.state:
            env = jm.state.get("env", {})
            if name in env:
                return env[name]
        raise AttributeError(f"'Environment' object has no attribute '{name}'")

    def __setattr__(self, name: str, value: Any) -> None:
        if name in ("model", "state"):
            super().__setattr__(name, value)
        else:
            if "state" not in self.__dict__:
                super().__setattr__("state", {})
            self.state[name] = value
            jm = getattr(self.__dict__.get("model"), "_jax_model", None)
            if jm is not None:
                jm.add_env_state(name, value)


class _Lazy:
    """An environment value computed when read (the materialised edge list)."""

    def __init__(self, fn):
        self.fn = fn

    def __call__(self):
        return self.fn()


class Grid:
    """2-D grid (jaxabm/agentpy.py:465-527).  ``shape`` is (width, height)."""

    def __init__(self, model: "Model", shape: Tuple[int, int], periodic: bool = False):
        self.model = model
        self.shape = shape
        self.periodic = periodic
        model.env.add_state("grid_shape", shape)
        model.env.add_state("grid_periodic", periodic)

    def position_agents(self, agents: AgentList, positions=None) -> None:
        """Random positions x = randint(PRNGKey(seed), (n,), 0, width), y =
        randint(split(PRNGKey(seed))[1], (n,), 0, height); a no-op until the
        collection has states."""
        n = len(agents)
        width, height = self.shape
        if positions is None:
            key = jr.PRNGKey(self.model.p.get("seed", 0))
            x = jr.randint(key, (n,), 0, width)
            sub = jr.split(key)[1]
            y = jr.randint(sub, (n,), 0, height)
            positions = torch.stack([x, y], dim=1)
        st = agents.collection.states
        if st is not None:
            new = dict(st)
            new["position"] = torch.as_tensor(positions, device=jr._device())
            agents.collection._states = new


    def _grid_array(self, values: torch.Tensor, dtype) -> torch.Tensor:
        t = torch.as_tensor(values, device=jr._device())
        if t.dim() != 2:
            raise ValueError("a Grid array is 2-D: (rows, cols)")
        return t.to(dtype).contiguous()

    def moore_sum(self, values) -> torch.Tensor:
        """Sum of ``values`` (a (rows, cols) array) over each cell's 8 Moore
        neighbours, wrapping when the grid is periodic (HIP, LDS-tiled;
        float32 result of a float64 sum in a fixed order)."""
        from ._lib import check, lib, ptr, stream_handle

        x = self._grid_array(values, torch.float32)
        out = torch.empty_like(x)
        check(lib().abm_grid_moore_sum_f32(x.shape[0], x.shape[1], int(bool(self.periodic)),
                                           ptr(x), ptr(out), stream_handle()), "grid_moore_sum")
        return out

    def moore_counts(self, cells) -> Tuple[torch.Tensor, torch.Tensor]:
        """For a (rows, cols) array of cell types (0 = empty): per cell, the
        number of Moore neighbours of its own non-zero type and the number of
        occupied Moore neighbours (int32), wrapping when periodic."""
        from ._lib import check, lib, ptr, stream_handle

        c = self._grid_array(cells, torch.int8)
        same = torch.empty(c.shape, dtype=torch.int32, device=c.device)
        occ = torch.empty(c.shape, dtype=torch.int32, device=c.device)
        check(lib().abm_grid_moore_counts_i8(c.shape[0], c.shape[1], int(bool(self.periodic)),
                                             ptr(c), ptr(same), ptr(occ), stream_handle()),
              "grid_moore_counts")
        return same, occ


class Network:
    """Agent network (jaxabm/agentpy.py:530-615)."""

    def __init__(self, model: "Model", directed: bool = False):
        self.model = model
        self.directed = directed
        self._src: List[int] = []
        self._dst: List[int] = []
        self._bulk: List[Tuple[np.ndarray, np.ndarray]] = []
        self._cache = None
        self._csr = None
        model.env.add_state("network_directed", directed)
        model.env.state["network_edges"] = _Lazy(self.edges)

    # -- construction
    def add_edge(self, from_agent: Union[Agent, int], to_agent: Union[Agent, int]) -> None:
        a = from_agent.id if isinstance(from_agent, Agent) else int(from_agent)
        b = to_agent.id if isinstance(to_agent, Agent) else int(to_agent)
        self._src.append(a)
        self._dst.append(b)
        if not self.directed and a != b:
            self._src.append(b)
            self._dst.append(a)
        self._invalidate()

    def add_edges(self, src, dst) -> None:
        """Bulk add_edge (both directions when undirected, in the same order
        add_edge would append them)."""
        s = np.asarray(src, np.int32).reshape(-1)
        d = np.asarray(dst, np.int32).reshape(-1)
        if not self.directed:
            # per edge: (a, b), then (b, a) unless a == b
            pairs = np.stack([np.stack([s, d], 1), np.stack([d, s], 1)], 1).reshape(-1, 2)
            keep = np.stack([np.ones(len(s), bool), s != d], 1).reshape(-1)
            flat = pairs[keep]
            s, d = flat[:, 0].copy(), flat[:, 1].copy()
        self._flush()
        self._bulk.append((s, d))
        self._invalidate()

    @classmethod
    def barabasi_albert(cls, model: "Model", n: int, m: int, seed: int = 0) -> "Network":
        """Undirected Barabasi-Albert graph from the engine's seeded generator."""
        from .network import ba_edges

        net = cls(model, directed=False)
        src, dst = ba_edges(n, m, seed)
        net.add_edges(src, dst)
        return net

    def _flush(self):
        if self._src:
            self._bulk.append((np.asarray(self._src, np.int32), np.asarray(self._dst, np.int32)))
            self._src, self._dst = [], []

    def _invalidate(self):
        self._cache = None
        self._csr = None

    def edges(self) -> torch.Tensor:
        """``env.network_edges``: (E, 2) int32 in insertion order."""
        if self._cache is None:
            self._flush()
            if self._bulk:
                s = np.concatenate([b[0] for b in self._bulk])
                d = np.concatenate([b[1] for b in self._bulk])
                self._bulk = [(s, d)]
            else:
                s = d = np.zeros(0, np.int32)
            self._cache = torch.from_numpy(np.stack([s, d], 1)).to(jr._device())
        return self._cache

    def _host_edges(self):
        self._flush()
        if not self._bulk:
            return np.zeros(0, np.int32), np.zeros(0, np.int32)
        s = np.concatenate([b[0] for b in self._bulk])
        d = np.concatenate([b[1] for b in self._bulk])
        self._bulk = [(s, d)]
        return s, d

    def _n_nodes(self, s, d) -> int:
        n = int(max(s.max(initial=-1), d.max(initial=-1))) + 1
        for lst in getattr(self.model, "_agent_lists", {}).values():
            n = max(n, len(lst))
        return n

    def csr(self):
        """CSR of the stored entries, built on the device (abm_csr_from_edges):
        rows ascending.  Undirected: row v holds the sorted unique ids of both
        directions of v's entries (what get_neighbors returns); directed: the
        targets of v's entries in insertion order, duplicates kept."""
        if self._csr is None:
            from .network import csr_from_edges

            s, d = self._host_edges()
            n = self._n_nodes(s, d)
            if self.directed:
                self._csr = csr_from_edges(n, s, d, insertion_order=True, device=jr._device())
            else:
                self._csr = csr_from_edges(n, s, d, symmetrize=True, dedup=True,
                                           device=jr._device())
        return self._csr

    # -- queries
    def get_neighbors(self, agent: Union[Agent, int]) -> torch.Tensor:
        """Directed: targets of the agent's edges in insertion order.
        Undirected: sorted unique ids over both directions.  One row of the
        device CSR (built once per change of the edge list)."""
        a = agent.id if isinstance(agent, Agent) else int(agent)
        return self.csr().row(a)

    def neighbor_sum(self, values: torch.Tensor) -> torch.Tensor:
        """out[v] = sum of values[u] over the CSR row of v (HIP segmented
        reduction)."""
        from .network import csr_gather_sum

        return csr_gather_sum(self.csr(), values)

    def neighbor_count(self, mask: torch.Tensor) -> torch.Tensor:
        """out[v] = number of neighbours u with mask[u] (int32)."""
        from .network import csr_count

        return csr_count(self.csr(), mask)


class Results:
    """Simulation results (jaxabm/agentpy.py:618-806)."""

    class VariableContainer:
        def __init__(self, data: Dict[str, Any]):
            self._data = data
            types = {k.split(".")[1] for k in data if k.startswith("agents.") and k.count(".") >= 2}
            for t in types:
                setattr(self, t, self.AgentContainer(data, t))

        class AgentContainer:
            def __init__(self, data: Dict[str, Any], agent_type: str):
                self._data = data
                self._agent_type = agent_type
                self._variables = set()
                for k in data:
                    if k.startswith(f"agents.{agent_type}."):
                        var = k.split(".")[-1]
                        self._variables.add(var)
                        setattr(self, var, self.VariableSeries(data[k], var))

            class VariableSeries:
                def __init__(self, values, name: str):
                    self._values = values
                    self._name = name

                def plot(self, ax=None, **kwargs):
                    plt = _pyplot()
                    if ax is None:
                        _, ax = plt.subplots()
                    vals = self._values
                    if len(vals) and hasattr(vals[0], "shape") and len(np.shape(vals[0])) > 0:
                        ax.plot(np.mean(np.array(vals), axis=1), **kwargs)
                        ax.set_ylabel(f"Mean {self._name}")
                    else:
                        ax.plot(vals, **kwargs)
                        ax.set_ylabel(self._name)
                    ax.set_xlabel("Time")
                    return ax

                def __getitem__(self, key: int):
                    return self._values[key]

                def __len__(self) -> int:
                    return len(self._values)

    def __init__(self, data: Dict[str, Any]):
        self._raw = data     # the device values as recorded (gradients flow through them)
        self._data = convert_to_numpy(data)
        self.variables = self.VariableContainer(self._data)

    def __getitem__(self, key: str):
        return self._data[key]

    def __setitem__(self, key: str, value) -> None:
        self._data[key] = value

    def __contains__(self, key: str) -> bool:
        return key in self._data

    def keys(self):
        return self._data.keys()

    def plot(self, variables: Optional[List[str]] = None, ax=None, **kwargs):
        plt = _pyplot()
        if ax is None:
            _, ax = plt.subplots()
        if variables is None:
            for k, v in self._data.items():
                if isinstance(v, list) and all(isinstance(x, (int, float, np.number)) for x in v):
                    ax.plot(v, label=k, **kwargs)
        else:
            for k in variables:
                if k in self._data:
                    ax.plot(self._data[k], label=k, **kwargs)
        ax.legend()
        ax.set_xlabel("Time")
        return ax

    def save(self, filename: str) -> None:
        import pickle

        with open(filename, "wb") as f:
            pickle.dump(self._data, f)

    @classmethod
    def load(cls, filename: str) -> "Results":
        """Load results saved by :meth:`save` (a pickle: load only files you
        wrote)."""
        import pickle

        with open(filename, "rb") as f:
            return cls(pickle.load(f))


def _pyplot():
    try:
        import matplotlib.pyplot as plt
    except ImportError as e:
        raise ImportError("plotting needs matplotlib, which is not installed") from e
    return plt


class Model:
    """Base class of façade models (jaxabm/agentpy.py:809-1152)."""

    def __init__(self, parameters: Optional[Dict[str, Any]] = None, seed: Optional[int] = None):
        self.p = parameters or {}
        self.seed = seed if seed is not None else self.p.get("seed", 0)
        self.steps = self.p.get("steps", 100)
        self.env = Environment(self)
        self._recorded_data: Dict[str, List[Any]] = {}
        self._agent_lists: Dict[str, AgentList] = {}
        self._jax_model: Optional[JaxModel] = None
        self._current_env_state: Dict[str, Any] = {}
        self._current_agent_states: Dict[str, Any] = {}
        self._running = False
        self._agent_instances: Dict[str, List[Agent]] = {}

    def setup(self) -> None:
        pass

    def step(self) -> None:
        pass

    def end(self) -> None:
        pass

    def update_state(self, env_state, agent_states, model_params, key) -> Dict[str, Any]:
        """Env-phase callback: store the current env and agents, run the user's
        ``step()``, return ``{**env_state, **self.env.state}``."""
        self._current_env_state = env_state.copy()
        self._current_agent_states = agent_states
        self.step()
        new = {**env_state}
        for name, value in self.env.state.items():
            new[name] = value() if isinstance(value, _Lazy) else value
        return new

    def compute_metrics(self, env_state, agent_states, model_params) -> Dict[str, Any]:
        return {}

    def add_agents(self, n: int, agent_class: Type[Agent], name: Optional[str] = None,
                   **kwargs) -> AgentList:
        lst = AgentList(self, n, agent_class, **kwargs)
        if name is None:
            name = agent_class.__name__.lower() + "s"
        self._agent_lists[name] = lst
        lst.name = name
        self._agent_instances[name] = []
        for i in range(n):
            a = agent_class()
            a.id = i
            a.model = self
            a.p = kwargs
            self._agent_instances[name].append(a)
        return lst

    def _update_agent_state(self, agent: Agent, new_state: Dict[str, Any]) -> None:
        for name, agents in self._agent_instances.items():
            if agent in agents:
                idx = agents.index(agent)
                jm = self._jax_model
                if self._running and jm is not None and jm.state:
                    coll_states = jm.state.get("agents", {}).get(name)
                    if coll_states is not None:
                        for k, v in new_state.items():
                            if k in coll_states:
                                t = coll_states[k].clone()
                                t[idx] = v
                                coll_states[k] = t
                agent._state.update(new_state)
                break

    def get_agent(self, collection_name: str, agent_id: int) -> Optional[Agent]:
        agents = self._agent_instances.get(collection_name)
        if agents is not None and 0 <= agent_id < len(agents):
            return agents[agent_id]
        return None

    def record(self, name: str, value: Any) -> None:
        self._recorded_data.setdefault(name, []).append(value)

    def run(self, steps: Optional[int] = None) -> Results:
        """setup() -> core model with this model's callbacks -> run -> end()
        -> Results of the metric history merged with record() lists."""
        if steps is not None:
            self.steps = steps
        self._running = True
        config = ModelConfig(steps=self.steps, collect_interval=1, seed=self.seed)
        self.setup()
        self._dynamic_state_update = type(self).step is not Model.step
        self._jax_model = JaxModel(params=self.p, config=config, update_state_fn=self.update_state,
                                   metrics_fn=self.compute_metrics)
        for name, lst in self._agent_lists.items():
            self._jax_model.add_agent_collection(name, lst.collection)
        for name, value in self.env.state.items():
            self._jax_model.add_env_state(name, value() if isinstance(value, _Lazy) else value)
        t0 = time.time()
        results = self._jax_model.run()
        t1 = time.time()
        self.end()
        self._running = False
        results.update(self._recorded_data)
        if self._jax_model.state and "agents" in self._jax_model.state:
            for name, states in self._jax_model.state["agents"].items():
                for var, values in states.items():
                    results[f"agents.{name}.{var}"] = values
        out = Results(results)
        print(f"Simulation executed in {format_time(t1 - t0)}")
        return out

    def batch_run(self, parameter_ranges: Dict[str, List[Any]],
                  repetitions: int = 1) -> Dict[tuple, List[Results]]:
        """Every parameter combination, ``repetitions`` runs with seeds
        seed + rep (jaxabm/agentpy.py:1116-1152)."""
        names = list(parameter_ranges.keys())
        combos = list(itertools.product(*parameter_ranges.values()))
        models = [type(self)({**self.p, **dict(zip(names, values))}, seed=self.seed + rep)
                  for values in combos for rep in range(repetitions)]
        from .programs.library import run_models, servable

        if len(models) > 1 and all(servable(m) for m in models):
            runs = run_models(models)   # one continuous batch, same results
        else:
            runs = [m.run() for m in models]
        return {tuple(values): runs[i * repetitions:(i + 1) * repetitions]
                for i, values in enumerate(combos)}


class Parameter:
    """A parameter with bounds for sampling (jaxabm/agentpy.py:1160-1205)."""

    def __init__(self, name: str, bounds: Tuple[float, float], distribution: str = "uniform"):
        self.name = name
        self.bounds = bounds
        self.distribution = distribution

    def sample(self, n: int = 1) -> np.ndarray:
        if self.distribution == "uniform":
            return np.random.uniform(self.bounds[0], self.bounds[1], size=n)
        if self.distribution == "normal":
            mean = (self.bounds[0] + self.bounds[1]) / 2
            std = (self.bounds[1] - self.bounds[0]) / 4
            return np.random.normal(mean, std, size=n)
        raise ValueError(f"Unknown distribution: {self.distribution}")


class Sample:
    def __init__(self, parameters: List[Parameter], n_samples: int = 10):
        self.parameters = parameters
        self.n_samples = n_samples
        self._samples = {p.name: p.sample(n_samples) for p in parameters}

    def __getitem__(self, index: int) -> Dict[str, float]:
        return {p.name: self._samples[p.name][index] for p in self.parameters}

    def __len__(self) -> int:
        return self.n_samples


class SensitivityAnalyzer:
    """Façade wrapper of :class:`jaxabm_amd.analysis.SensitivityAnalysis`."""

    def __init__(self, model_class: Type[Model], parameters: List[Parameter], n_samples: int = 10,
                 metrics: List[str] = None):
        from .analysis import SensitivityAnalysis

        self.model_class = model_class
        self.parameters = parameters
        self.n_samples = n_samples
        self.metrics = metrics or []
        self.sample = Sample(parameters, n_samples)

        def model_factory(params=None, config=None):
            return self.model_class(params)

        self.analysis = SensitivityAnalysis(
            model_factory=model_factory, param_ranges={p.name: p.bounds for p in parameters},
            metrics_of_interest=self.metrics, num_samples=n_samples)

    def run(self) -> Dict[str, Any]:
        out = self.analysis.run()
        return out._data if hasattr(out, "_data") else out

    def calculate_sensitivity(self, method: str = "sobol") -> Dict[str, Any]:
        if method == "sobol":
            return self.analysis.sobol_indices()
        if method == "morris":
            return self.analysis.morris_indices()
        raise ValueError(f"Unknown method: {method}")

    def plot(self, metric: Optional[str] = None, ax=None, **kwargs):
        return self.analysis.plot(metric, ax, **kwargs)


class ModelCalibrator:
    """Façade wrapper of :class:`jaxabm_amd.analysis.ModelCalibrator`."""

    def __init__(self, model_class: Type[Model], parameters: List[Parameter],
                 target_metrics: Dict[str, float], metrics_weights: Dict[str, float] = None,
                 learning_rate: float = 0.01, max_iterations: int = 20, method: str = "gradient"):
        from .analysis import ModelCalibrator as _Cal

        self.model_class = model_class
        self.parameters = parameters
        self.target_metrics = target_metrics
        self.metrics_weights = metrics_weights or {m: 1.0 for m in target_metrics}
        self.learning_rate = learning_rate
        self.max_iterations = max_iterations
        self.method = method

        def model_factory(params=None, config=None):
            return self.model_class(params)

        self.calibrator = _Cal(
            model_factory=model_factory,
            initial_params={p.name: (p.bounds[0] + p.bounds[1]) / 2 for p in parameters},
            target_metrics=target_metrics, metrics_weights=metrics_weights,
            param_bounds={p.name: tuple(p.bounds) for p in parameters},
            learning_rate=learning_rate, max_iterations=max_iterations,
            method="adam" if method == "gradient" else method)

    def run(self) -> Dict[str, float]:
        return self.calibrator.calibrate()

    def plot_progress(self, ax=None, **kwargs):
        plt = _pyplot()
        if ax is None:
            _, ax = plt.subplots()
        if self.calibrator.loss_history:
            ax.plot(self.calibrator.loss_history, **kwargs)
            ax.set_xlabel("Iteration")
            ax.set_ylabel("Loss")
        return ax


__all__ = ["Agent", "AgentWrapper", "AgentTypeWrapper", "AgentList", "Environment", "Grid",
           "Network", "Model", "Results", "Parameter", "Sample", "SensitivityAnalyzer",
           "ModelCalibrator"]
