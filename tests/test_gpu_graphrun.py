"""GPU: generic model steps replayed as a hipGraph (Model.run(graph=True))
against the eager loop of the same model (jaxabm/model.py:146-262).

Every replay is checked: the history holds each step's metrics, and the
replayed run must equal the eager run bit for bit at every step, with eager
work and synchronisations between replays in one case.  The large model's
full reductions span many workgroups, whose semaphores torch clears with a
memset: the capture holds memset nodes that abm_graph_sanitize replaces by
fill kernels."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def J():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import jaxabm_amd

    return jaxabm_amd


class Dummy:
    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        return {"value": jr.uniform(key, minval=0.0, maxval=10.0)}

    def update(self, state, model_state, cfg, key):
        return {"value": state["value"] + model_state["env"].get("increment", 1.0)}


def counter_fn(env, agents, params, key):
    new = dict(env)
    new["counter"] = env.get("counter", 0) + 1
    return new


def core_metrics(env, agents, params):
    from jaxabm_amd import numpy as jnp

    return {"total_value": jnp.sum(agents["c"]["value"]), "step_counter": env.get("counter", 0)}


class Trader:
    """Per-agent position walk and a coin-flip wealth change, with a constant
    output and the time step in the arithmetic."""

    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        return {"x": jr.uniform(k1), "w": jr.randint(k2, (), 0, 6), "flag": 1}

    def update(self, state, model_state, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        x = jnp.mod(state["x"] + jr.normal(k1) * 0.01, 1.0)
        give = jr.uniform(k2) < 0.5
        w = jnp.where(give & (state["w"] > 0), state["w"] - 1, state["w"] + 1)
        w = w + (model_state["time_step"] % 2)
        return {"x": x, "w": w, "flag": 1}


def trader_env(env, agents, params, key):
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    new = dict(env)
    new["mean_x"] = jnp.mean(agents["t"]["x"])
    new["shock"] = jr.uniform(key) * params["scale"]
    new["t"] = env.get("t", 0) + 1
    return new


def trader_metrics(env, agents, params):
    from jaxabm_amd import numpy as jnp

    a = agents["t"]
    return {"mean_x": env["mean_x"], "total_w": jnp.sum(a["w"]), "max_x": jnp.max(a["x"]),
            "rich": jnp.sum(a["w"] > 3), "shock": env["shock"], "t": env["t"]}


def make_core(J, seed=42):
    m = J.JaxModel(params={}, config=J.ModelConfig(steps=20, seed=seed),
                   update_state_fn=counter_fn, metrics_fn=core_metrics)
    m.add_agent_collection("c", J.AgentCollection(Dummy(), 10))
    m.add_env_state("increment", 1.0)
    return m


def make_trader(J, n=262144, seed=7, interval=1, metrics=trader_metrics, env_fn=trader_env):
    m = J.JaxModel(params={"scale": 0.5},
                   config=J.ModelConfig(steps=30, seed=seed, collect_interval=interval),
                   update_state_fn=env_fn, metrics_fn=metrics)
    m.add_agent_collection("t", J.AgentCollection(Trader(), n))
    return m


def _host(v):
    import torch

    if isinstance(v, torch.Tensor):
        return v.detach().cpu().numpy()
    return np.asarray(v)


def assert_same_run(a, b, ra, rb):
    assert list(ra) == list(rb)
    for k in ra:
        assert len(ra[k]) == len(rb[k]), k
        for i, (x, y) in enumerate(zip(ra[k], rb[k])):
            assert type(x) is type(y) or (hasattr(x, "shape") and hasattr(y, "shape")), (k, i)
            np.testing.assert_array_equal(_host(x), _host(y), err_msg=f"{k} step {i}")
    for name in a.agent_collections:
        sa, sb = a.agent_collections[name].states, b.agent_collections[name].states
        assert set(sa) == set(sb)
        for k in sa:
            assert sa[k].dtype == sb[k].dtype
            np.testing.assert_array_equal(_host(sa[k]), _host(sb[k]), err_msg=f"{name}.{k}")
    assert set(a._env_state) == set(b._env_state)
    for k in a._env_state:
        x, y = a._env_state[k], b._env_state[k]
        assert type(x) is type(y) or (hasattr(x, "shape") and hasattr(y, "shape")), k
        np.testing.assert_array_equal(_host(x), _host(y), err_msg=f"env {k}")
    np.testing.assert_array_equal(a._rng, b._rng)
    assert a._time_step == b._time_step


def test_core_model_graph_matches_eager(J, capsys):
    a, b = make_core(J), make_core(J)
    ra = a.run(20)
    rb = b.run(20, graph=True)
    info = b.last_run_info
    print(info)
    assert info["mode"] == "graph" and info["steps_done"] == 20
    assert_same_run(a, b, ra, rb)
    assert b._env_state["counter"] == 20 and isinstance(b._env_state["counter"], int)
    assert rb["step_counter"] == list(range(1, 21))
    # a second run continues from the replayed state
    assert_same_run(a, b, a.run(7), b.run(7, graph=True))


def test_large_model_graph_matches_eager_every_replay(J):
    """262,144 agents: draws under vmap, multi-workgroup reductions (memset
    nodes replaced), env callback with its own key; eager work and a
    synchronisation between replays."""
    import torch

    from jaxabm_amd.graphrun import run_graph

    a, b = make_trader(J), make_trader(J)
    ra = a.run(30)
    b.initialize()
    b._history = []
    scratch = torch.rand(1 << 22, device="cuda")

    def between():
        scratch.mul_(1.0001).sum()
        torch.cuda.synchronize()

    info = run_graph(b, 30, between=between)
    print(info)
    assert info["mode"] == "graph" and info["steps_done"] == 30
    assert info["kernel_nodes"] > 0
    assert info["memsets_replaced"] > 0, "the reductions should have recorded memset nodes"
    out = {"step": [h["time_step"] for h in b._history]}
    for k in b._history[0]["metrics"]:
        out[k] = [h["metrics"][k] for h in b._history]
    ra2 = {k: v for k, v in ra.items()}
    assert_same_run(a, b, ra2, out)
    assert out["t"] == list(range(1, 31))


def test_batched_replays_match_eager(J, monkeypatch):
    """A run long enough for graphs of several step bodies (here 4 a launch
    from 40 steps, 43 steps: ten 4-step replays and one 1-step replay, with
    eager work and a synchronisation between replays) equals the eager loop,
    history and collect_interval included, and a continued run too."""
    import torch

    from jaxabm_amd import graphrun

    monkeypatch.setattr(graphrun, "STEPS_PER_LAUNCH", 4)
    monkeypatch.setattr(graphrun, "BATCH_MIN_STEPS", 40)
    a, b = make_trader(J, n=20_000, interval=2), make_trader(J, n=20_000, interval=2)
    ra = a.run(43)
    b.initialize()
    b._history = []
    scratch = torch.rand(1 << 20, device="cuda")
    calls = []

    def between():
        calls.append(1)
        scratch.mul_(1.0001).sum()
        torch.cuda.synchronize()

    info = graphrun.run_graph(b, 43, between=between)
    assert info["mode"] == "graph" and info["steps_per_launch"] == 4 and info["steps_done"] == 43
    assert len(calls) == 10       # 11 replays
    out = {"step": [h["time_step"] for h in b._history]}
    for k in b._history[0]["metrics"]:
        out[k] = [h["metrics"][k] for h in b._history]
    assert_same_run(a, b, dict(ra), out)
    assert_same_run(a, b, a.run(45), b.run(45, graph=True))
    assert b.last_run_info["steps_per_launch"] == 4


def test_collect_interval_and_continuation(J):
    a, b = make_trader(J, n=5000, interval=3), make_trader(J, n=5000, interval=3)
    assert_same_run(a, b, a.run(10), b.run(10, graph=True))
    assert b.last_run_info["mode"] == "graph"
    assert_same_run(a, b, a.run(8), b.run(8, graph=True))


def test_host_read_in_step_runs_eagerly(J):
    """A metric read on the host (float()) cannot be replayed: step 2 (run
    with synchronising calls raising) stops the replay and the run continues
    eagerly, with the same results."""
    from jaxabm_amd import numpy as jnp

    def host_metrics(env, agents, params):
        return {"mean_x": float(jnp.mean(agents["t"]["x"]))}

    a = make_trader(J, n=4096, metrics=host_metrics)
    b = make_trader(J, n=4096, metrics=host_metrics)
    ra, rb = a.run(6), b.run(6, graph=True)
    info = b.last_run_info
    assert info["mode"] == "eager" and info["steps_done"] == 1 and "step 2" in info["reason"]
    assert_same_run(a, b, ra, rb)


def test_python_value_changing_between_steps_runs_eagerly_after_step_1(J):
    """An env value computed from Python-side state changes every step: it
    would be a constant of the graph, so step 2's result is discarded and
    steps 2..n run eagerly.  Step 2's Python code ran once more (the counter
    below is one ahead of the eager run's); the device state is that of the
    eager run."""

    def make():
        calls = {"n": 0}

        def env_fn(env, agents, params, key):
            calls["n"] += 1
            new = trader_env(env, agents, params, key)
            new["label"] = "odd" if calls["n"] % 2 else "even"
            return new

        return make_trader(J, n=4096, env_fn=env_fn), calls

    (a, ca), (b, cb) = make(), make()
    ra, rb = a.run(5), b.run(5, graph=True)
    info = b.last_run_info
    assert info["mode"] == "eager" and info["steps_done"] == 1 and "step 2" in info["reason"]
    assert (ca["n"], cb["n"]) == (5, 6)
    assert (a._env_state.pop("label"), b._env_state.pop("label")) == ("odd", "even")
    assert_same_run(a, b, ra, rb)


def test_differentiable_run_stays_eager(J, capsys):
    """A run whose env value requires grad is not replayed (a graph would cut
    it from autograd): it runs eagerly and the gradient reaches the value."""
    import torch

    from jaxabm_amd import numpy as jnp

    inc = torch.tensor(0.5, device="cuda", requires_grad=True)
    m = J.JaxModel(config=J.ModelConfig(seed=1),
                   metrics_fn=lambda env, agents, p: {"m": jnp.mean(agents["d"]["value"])})
    m.add_agent_collection("d", J.AgentCollection(Dummy(), 1000))
    m.add_env_state("increment", inc)
    out = m.run(6, graph=True)
    assert m.last_run_info["mode"] == "eager" and "grad" in m.last_run_info["reason"]
    (g,) = torch.autograd.grad(out["m"][-1], [inc])
    assert float(g) == pytest.approx(6.0, rel=1e-6)


def test_copy_items_paths(J):
    """abm_copy_items: 16-byte and byte paths, history rows at row[0] *
    row_stride, rows at or past row_limit skipped; abm_step_counters."""
    import ctypes as C

    import torch

    from jaxabm_amd._lib import CopyItem, check, lib, stream_handle

    a = torch.arange(1000, dtype=torch.float32, device="cuda")          # 4000 B, aligned
    b = torch.arange(7, dtype=torch.uint8, device="cuda") + 1           # 7 B, byte path
    da, db = torch.zeros_like(a), torch.zeros(7, dtype=torch.uint8, device="cuda")
    hist = torch.zeros((5, 3), dtype=torch.float32, device="cuda")
    v = torch.tensor([1.5, 2.5, 3.5], device="cuda")
    row = torch.tensor([2], dtype=torch.int64, device="cuda")
    ts = torch.zeros((), dtype=torch.int32, device="cuda")
    items = (CopyItem * 3)(CopyItem(a.data_ptr(), da.data_ptr(), 4000, 0, 0),
                           CopyItem(b.data_ptr(), db.data_ptr(), 7, 0, 0),
                           CopyItem(v.data_ptr(), hist.data_ptr(), 12, 12, 5))
    check(lib().abm_copy_items(items, 3, row.data_ptr(), stream_handle()), "copy_items")
    check(lib().abm_step_counters(row.data_ptr(), ts.data_ptr(), stream_handle()), "counters")
    torch.cuda.synchronize()
    assert torch.equal(da, a) and torch.equal(db, b)
    want = torch.zeros((5, 3))
    want[2] = v.cpu()
    assert torch.equal(hist.cpu(), want)
    assert int(row.item()) == 3 and int(ts.item()) == 1
    row.fill_(5)                                   # at row_limit: not written
    check(lib().abm_copy_items(items, 3, row.data_ptr(), stream_handle()), "copy_items")
    torch.cuda.synchronize()
    assert torch.equal(hist.cpu(), want)
    with pytest.raises(Exception):
        lib_items = (CopyItem * 1)(CopyItem(v.data_ptr(), hist.data_ptr(), 12, 12, 0))
        check(lib().abm_copy_items(lib_items, 1, row.data_ptr(), stream_handle()), "bad")


def test_commit_step_tail(J):
    """abm_commit_step: the copies, then (once, by the last block) row and
    time step advance, the counter returns to 0 and the next row of the key
    table lands in the key buffer; past the table's last row no keys are
    copied."""
    import torch

    from jaxabm_amd._lib import CopyItem, check, lib, stream_handle

    a = torch.arange(300_000, dtype=torch.float32, device="cuda")     # many blocks
    da = torch.zeros_like(a)
    table = torch.arange(3 * 4, dtype=torch.int32, device="cuda").reshape(3, 2, 2)
    keys = table[0].clone()
    row = torch.zeros(1, dtype=torch.int64, device="cuda")
    ts = torch.zeros((), dtype=torch.int32, device="cuda")
    done = torch.zeros(1, dtype=torch.int32, device="cuda")
    items = (CopyItem * 1)(CopyItem(a.data_ptr(), da.data_ptr(), a.numel() * 4, 0, 0))
    for step in range(1, 4):
        check(lib().abm_commit_step(items, 1, row.data_ptr(), ts.data_ptr(), done.data_ptr(),
                                    table.data_ptr(), 3, 4, keys.data_ptr(), stream_handle()),
              "commit_step")
        torch.cuda.synchronize()
        assert int(row.item()) == step and int(ts.item()) == step and int(done.item()) == 0
        assert torch.equal(keys, table[min(step, 2)])
    assert torch.equal(da, a)
    empty = (CopyItem * 1)()
    check(lib().abm_commit_step(empty, 0, row.data_ptr(), ts.data_ptr(), done.data_ptr(),
                                table.data_ptr(), 3, 4, keys.data_ptr(), stream_handle()), "none")
    torch.cuda.synchronize()
    assert int(row.item()) == 4 and int(done.item()) == 0


class Trader:
    """The generic workload's agent (a normal-draw walk of x, a coin-flip
    wealth change), as the oracle restates it (orc_trader_*)."""

    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        return {"x": jr.uniform(k1), "w": jr.randint(k2, (), 0, 6)}

    def update(self, state, model_state, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        x = jnp.mod(state["x"] + jr.normal(k1) * 0.01, 1.0)
        give = jr.uniform(k2) < 0.5
        return {"x": x, "w": jnp.where(give & (state["w"] > 0), state["w"] - 1, state["w"] + 1)}


@pytest.mark.parametrize("fused,graph", [(True, True), (False, False)])
def test_generic_trader_matches_oracle(J, capsys, fused, graph):
    """The generic workload's model (10,000 agents, 200 steps, an env phase)
    against the oracle's restatement, bit for bit (wealth, positions and the
    final key), in the graph-replayed fused run and the eager vmapped one."""
    import oracle as O
    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp

    n, steps = 10_000, 200
    fuse.enable(fused)
    try:
        m = J.JaxModel(params={}, config=J.ModelConfig(steps=steps, seed=0),
                       update_state_fn=lambda env, a, p, k: {**env, "mean_x": jnp.mean(a["t"]["x"])},
                       metrics_fn=lambda env, a, p: {"total_w": jnp.sum(a["t"]["w"])})
        m.add_agent_collection("t", J.AgentCollection(Trader(), n))
        res = m.run(steps, graph=graph)
    finally:
        fuse.enable(True)
    assert m.last_run_info["mode"] == ("graph" if graph else "eager")
    x, w, rng = O.trader_init(n, 0, O.MODE_PARTITIONABLE)
    O.trader_run(x, w, rng, steps, O.MODE_PARTITIONABLE)
    st = m.agent_collections["t"].states
    np.testing.assert_array_equal(st["w"].cpu().numpy(), w)
    np.testing.assert_array_equal(st["x"].cpu().numpy(), x)
    assert int(res["total_w"][-1]) == int(w.sum())
    np.testing.assert_array_equal(np.asarray(m._rng, np.uint32), rng)


def test_env_callback_that_updates_its_env_in_place(J):
    """An env callback that writes into the dict it receives and returns that
    same dict (env['p'] = f(env['p']); return env) replays as the eager loop
    does: every step advances from the previous step's env."""
    import torch

    from jaxabm_amd import numpy as jnp

    def inplace(env, agents, params, key):
        env["p"] = env["p"] * 1.5 + jnp.mean(agents["c"]["value"])
        env["q"] = env["q"] + 1
        return env

    def metrics(env, agents, params):
        return {"p": env["p"], "q": env["q"]}

    out = []
    for graph in (False, True):
        m = J.JaxModel(params={}, config=J.ModelConfig(steps=30, seed=3),
                       update_state_fn=inplace, metrics_fn=metrics)
        m.add_agent_collection("c", J.AgentCollection(Dummy(), 64))
        m.add_env_state("p", torch.tensor(0.5, device="cuda"))
        m.add_env_state("q", torch.tensor(0, dtype=torch.int32, device="cuda"))
        m.add_env_state("increment", 0.25)
        res = m.run(30, graph=graph)
        assert m.last_run_info["mode"] == ("graph" if graph else "eager")
        out.append(([float(v) for v in res["p"]], [int(v) for v in res["q"]]))
    assert out[0] == out[1]
    assert out[0][1] == list(range(1, 31))
