"""Host-side behaviour of the façade and core API (no GPU work): the
reference's structural unit tests (tests/unit/test_agent.py:141-155,
tests/unit/test_model.py:241-300, tests/unit/test_api.py:22-150,
tests/unit/test_utils.py) where they need no device."""
import pickle

import numpy as np
import pytest
import torch

import jaxabm_amd as J
from jaxabm_amd.agentpy import Agent, AgentTypeWrapper, Results
from jaxabm_amd.utils import (convert_to_numpy, format_time, is_valid_params, mean_over_runs,
                              standardize_metrics)


def test_model_config():
    c = J.ModelConfig(seed=42, steps=10, track_history=True, collect_interval=2)
    assert (c.seed, c.steps, c.track_history, c.collect_interval) == (42, 10, True, 2)
    d = J.ModelConfig()
    assert (d.seed, d.steps, d.track_history, d.collect_interval) == (0, 100, True, 1)


class Dummy:
    def init_state(self, model_config, key):
        return {"value": 0.0}

    def update(self, state, model_state, model_config, key):
        return state


def test_agent_collection_errors():
    for bad in (0, -3, 2.5, "4"):
        with pytest.raises(ValueError, match="num_agents must be a positive integer"):
            J.AgentCollection(Dummy(), bad)
    c = J.AgentCollection(Dummy(), 4)
    with pytest.raises(TypeError, match="model_config must be a ModelConfig instance."):
        c.init(None, {"seed": 0})
    with pytest.raises(ValueError, match="Agent collection not initialized"):
        c.update({}, None, J.ModelConfig())


def test_core_model_errors_and_state():
    m = J.JaxModel(config=J.ModelConfig(steps=3))
    with pytest.raises(ValueError, match="No agent collections added to model"):
        m.initialize()
    m.add_agent_collection("a", J.AgentCollection(Dummy(), 2))
    with pytest.raises(RuntimeError, match="initialized"):
        m.step()
    m2 = J.JaxModel(config=J.ModelConfig())
    m2.add_env_state("price_level", 1.0)
    assert m2.state == {"env": {"price_level": 1.0}}
    assert m2._time_step == 0 and m2._history == [] and not m2._is_initialized


def test_agent_defaults_and_attributes():
    a = Agent()
    assert a.id is None and a.model is None and a.p == {} and a._state == {}
    assert a.setup() == {}
    a._state = {"x": 1, "y": 2}
    assert a.step() == {"x": 1, "y": 2}
    assert a.x == 1
    a.x = 5
    assert a._state["x"] == 5
    with pytest.raises(AttributeError):
        _ = a.nope


class Walker(Agent):
    def setup(self):
        return {"x": 0, "y": 0, "energy": 100}

    def step(self, model_state=None):
        s = self._state.copy()
        s["x"] += 1
        s["y"] += 1
        s["energy"] -= 1
        return s


def test_agent_wrapper():
    w = AgentTypeWrapper(Walker, {"max_energy": 100, "speed": 1})
    assert w.agent_class is Walker and w.params == {"max_energy": 100, "speed": 1}
    assert isinstance(w.agent_instance, Walker) and w.agent_instance.p == w.params
    assert AgentTypeWrapper(Walker).params == {} and AgentTypeWrapper(Walker).agent_instance.p == {}
    cfg = J.ModelConfig(seed=42)
    assert w.init_state(cfg, None) == {"x": 0, "y": 0, "energy": 100}
    assert w.update({"x": 5, "y": 10, "energy": 50}, {"g": 1}, cfg, None) == {
        "x": 6, "y": 11, "energy": 49}

    class Bad(Agent):
        def setup(self):
            return "invalid"

        def step(self, model_state=None):
            return "invalid"

    class NoneAgent(Agent):
        def setup(self):
            return None

        def step(self, model_state=None):
            return None

    with pytest.raises(ValueError, match=r"Agent.setup\(\) must return a dictionary"):
        AgentTypeWrapper(Bad).init_state(cfg, None)
    with pytest.raises(ValueError, match=r"Agent.step\(\) must return a dictionary"):
        AgentTypeWrapper(Bad).update({"x": 1}, {}, cfg, None)
    assert AgentTypeWrapper(NoneAgent).init_state(cfg, None) == {}
    assert AgentTypeWrapper(NoneAgent).update({"x": 1}, {}, cfg, None) == {"x": 1}


def test_utils():
    data = {"a": torch.arange(3), "b": [torch.ones(2), (torch.zeros(1), 5)], "c": "s"}
    out = convert_to_numpy(data)
    assert isinstance(out["a"], np.ndarray) and out["a"].tolist() == [0, 1, 2]
    assert isinstance(out["b"][0], np.ndarray) and isinstance(out["b"][1], tuple)
    assert out["b"][1][1] == 5 and out["c"] == "s"
    assert format_time(12.345) == "12.35s"
    assert format_time(125.5) == "2m 5.50s"
    assert format_time(3725.0) == "1h 2m 5.00s"
    assert is_valid_params({"a": 1, "b": 2}, ["a"]) and not is_valid_params({"a": 1}, ["a", "b"])
    assert mean_over_runs([{"x": [1.0, 2.0]}, {"x": [3.0, 4.0]}]) == {"x": [2.0, 3.0]}
    assert mean_over_runs([]) == {}
    assert standardize_metrics({"a": np.float32(1.5), "b": 2, "c": "x"}) == {"a": 1.5, "b": 2.0}


def test_results_container_and_roundtrip(tmp_path):
    r = Results({"step": [1, 2, 3], "mean": [0.1, 0.2, 0.3],
                 "agents.walkers.wealth": [np.ones(4), 2 * np.ones(4), 3 * np.ones(4)],
                 "t": torch.tensor([1.0, 2.0])})
    assert isinstance(r._data["t"], np.ndarray)
    series = r.variables.walkers.wealth
    assert len(series) == 3 and series[2].tolist() == [3.0] * 4
    assert "mean" in r and r["step"] == [1, 2, 3]
    path = tmp_path / "res.pkl"
    r.save(str(path))
    back = Results.load(str(path))
    assert back["mean"] == [0.1, 0.2, 0.3]
    with open(path, "rb") as f:
        assert pickle.load(f)["step"] == [1, 2, 3]
