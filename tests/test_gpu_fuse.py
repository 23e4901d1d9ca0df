"""GPU: agent functions run as one generated kernel (jaxabm_amd.fuse) give
the vmapped torch program's results bit for bit, NaN results as NaN
(jaxabm/agent.py:92-177:
the reference's vmap of init_state / update), over the covered operations,
special values (signed zeros, NaN, negative operands of // and %), env
scalars of each kind, the time step and every draw, in both stream modes."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def J():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import jaxabm_amd

    return jaxabm_amd


class Arith:
    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        k1, k2, k3 = jr.split(key, 3)
        return {"x": jr.uniform(k1, minval=-3.0, maxval=3.0), "w": jr.randint(k2, (), -7, 8),
                "b": jr.bernoulli(k3, 0.4), "z": 0.0}

    def update(self, s, ms, cfg, key):
        from jaxabm_amd import numpy as jnp

        env = ms["env"]
        x, w, b = s["x"], s["w"], s["b"]
        sz = x * 0.0                                   # signed zeros
        nan = jnp.log(x)                               # NaN where x < 0
        y = (x + env["shift"]) * env["scale"] - w / 3 + (x ** 2) / 7.0
        q = w // 3 + w % 4 + (w * env["k"]) // -2 + jnp.mod(w, -3)
        fq = x // 0.7 + jnp.mod(x, 1.3) + jnp.mod(x, -0.9)
        m = jnp.minimum(nan, sz) + jnp.maximum(sz, -sz) * 2 + jnp.minimum(x, y)
        c = jnp.clip(y, -1.0, 1.5) + jnp.clip(nan, 0.0, 1.0)
        t = (jnp.exp(x / 4) + jnp.log(jnp.abs(x) + 1.0) + jnp.sqrt(jnp.abs(y)) + jnp.sin(x)
             - jnp.cos(y) + jnp.tanh(x * 2) + jnp.floor(y) - jnp.ceil(x) + jnp.round(x * 3)
             + jnp.sign(sz + x))
        flag = (b & (x > env["thr"])) | ((w < 0) ^ env["on"])
        k = jnp.where(flag, w + ms["time_step"], -w) + jnp.asarray(y * 10, dtype=jnp.int32)
        # torch divides by a Python scalar through its reciprocal
        d = (x / 0.7 + 3.0 / (x + 5.0) + x // env["scale"] + w / 3 + x / env["delta"]
             + x // -1.3)
        return {"x": jnp.where(b, y, t) + env["delta"], "w": q + k, "b": ~flag,
                "z": m + c + fq + s["z"] * 0.5 + d}


class Draws:
    def init_state(self, cfg, key):
        from jaxabm_amd import random as jr

        return {"x": jr.normal(key), "i": jr.bits(key)}

    def update(self, s, ms, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2, k3, k4 = jr.split(key, 4)
        a, b = jr.split(k1)
        lo = jnp.minimum(s["x"], 0.0)
        return {"x": jr.uniform(a, minval=lo, maxval=lo + 2.0) + jr.normal(b) * 0.5,
                "i": jr.bits(k2) ^ s["i"],
                "r": jr.randint(k3, (), -20, 11) + jr.choice(k4, 5),
                "p": jr.bernoulli(k4, jnp.abs(s["x"]) / 4.0), "q": jr.bernoulli(b, 0.3),
                "u": jr.uniform(k2, minval=0.5, maxval=3.25)}


class Walker:
    """A bouncing walker in the style of the reference's basic example
    (examples/basic_example.py:20-70): per-agent 2-vectors, element reads, an
    env bounds vector, jnp.array of per-agent values, vector draws."""

    def init_state(self, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2 = jr.split(key)
        return {"position": jr.uniform(k1, (2,)), "velocity": jr.normal(k2, (2,)) * 0.05,
                "color": 0, "steps_taken": 0, "home": jnp.array([0.5, 0.25])}

    def update(self, s, ms, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        position, velocity = s["position"], s["velocity"]
        bounds = ms["env"]["bounds"]
        new_position = position + velocity
        x_bounce = (new_position[0] <= bounds[0]) | (new_position[0] >= bounds[1])
        y_bounce = (new_position[1] <= bounds[0]) | (new_position[1] >= bounds[1])
        new_velocity = velocity * jnp.array([1 - 2 * x_bounce, 1 - 2 * y_bounce])
        new_position = jnp.clip(new_position, bounds[0], bounds[1])
        k1, k2, k3 = jr.split(key, 3)
        jitter = jr.uniform(k1, (2,), minval=-0.01, maxval=0.01)
        lo = jnp.minimum(new_position, 0.5)
        return {"position": new_position + jitter * (s["home"] - new_position) ** 2,
                "velocity": jnp.where(x_bounce | y_bounce, -new_velocity, new_velocity) / 1.5,
                "color": jnp.where(jnp.logical_or(x_bounce, y_bounce), 1 - s["color"], s["color"]),
                "steps_taken": s["steps_taken"] + 1, "home": s["home"],
                "r": jr.randint(k2, (3,), -4, 9), "u": jr.uniform(k3, (2,), minval=lo,
                                                                  maxval=lo + 1.0)}


def run(J, agent, n, steps, mode, fused, env):
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import random as jr

    fuse.enable(fused)
    jr.set_threefry_partitionable(mode == O.MODE_PARTITIONABLE)
    try:
        m = J.JaxModel(config=J.ModelConfig(seed=3))
        m.add_agent_collection("a", J.AgentCollection(agent, n))
        for k, v in env.items():
            m.add_env_state(k, v)
        m.initialize()
        init = {k: v.clone() for k, v in m.agent_collections["a"].states.items()}
        for _ in range(steps):
            m.step()
        torch.cuda.synchronize()
        return init, {k: v.clone() for k, v in m.agent_collections["a"].states.items()}
    finally:
        fuse.enable(True)
        jr.set_threefry_partitionable(True)


def bits(t):
    """Bit patterns; every NaN compares as one NaN (positions must agree).
    torch subtracts as a + (-1) * b with a run-time multiply, the kernel as
    a - b, whose negate modifier flips the sign bit of a NaN operand: every
    non-NaN result is the same, a NaN result's sign bit may differ."""
    import torch

    if t.dtype == torch.float32:
        b = t.view(torch.int32).cpu().numpy().copy()
        b[np.isnan(t.cpu().numpy())] = 0x7FC00000
        return b
    return t.cpu().numpy()


@pytest.mark.parametrize("mode", [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE])
@pytest.mark.parametrize("agent,n", [(Arith, 1000), (Arith, 257), (Draws, 1000), (Draws, 3),
                                     (Walker, 5000), (Walker, 1)])
def test_fused_equals_vmapped(J, agent, n, mode):
    import torch

    from jaxabm_amd import fuse

    env = {"shift": 0.25, "scale": 1.5, "k": 3, "thr": 0.1, "on": True,
           "delta": torch.tensor(0.125, device="cuda"),
           "bounds": torch.tensor([0.0, 1.0], device="cuda")}
    before = fuse.stats()["launches"]
    i_ref, ref = run(J, agent(), n, 4, mode, False, env)
    mid = fuse.stats()["launches"]
    i_got, got = run(J, agent(), n, 4, mode, True, env)
    after = fuse.stats()["launches"]
    assert mid == before, "the vmapped run must not use the fused path"
    assert after - mid == 5, (after - mid, fuse.refused())     # init + 4 updates fused
    for init_or_end, (a, b) in {"init": (i_ref, i_got), "end": (ref, got)}.items():
        assert set(a) == set(b)
        for k in a:
            assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, (init_or_end, k)
            np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=f"{init_or_end} {k}")
    if agent is Arith:
        x = ref["z"]
        assert bool(torch.isnan(x).any()), "the case should exercise NaN"


def test_refused_function_runs_vmapped(J):
    """A difference of a per-agent vector is not covered: the update runs
    the vmapped path (and the refusal is remembered); the init is fused."""
    from jaxabm_amd import fuse
    from jaxabm_amd import random as jr

    class Vec:
        def init_state(self, cfg, key):
            return {"p": jr.uniform(key, (2,))}

        def update(self, s, ms, cfg, key):
            from jaxabm_amd import numpy as jnp

            return {"p": s["p"] + jnp.diff(s["p"])}

    before = fuse.stats()["launches"]
    c = J.AgentCollection(Vec(), 50)
    c.init(jr.PRNGKey(1), J.ModelConfig())
    p0 = c.states["p"].clone()
    c.update({"env": {}}, jr.PRNGKey(2), J.ModelConfig())
    assert c.states["p"].shape == (50, 2)
    assert fuse.stats()["launches"] == before + 1
    assert any("diff" in r for r in fuse.refused().values())
    import torch

    assert torch.equal(c.states["p"], p0 + (p0[:, 1:] - p0[:, :1]))


OPS = {
    "add_pyf": lambda s, m: s["x"] + m["env"]["shift"],
    "mul_pyf": lambda s, m: (s["x"] + m["env"]["shift"]) * m["env"]["scale"],
    "div_s": lambda s, m: s["x"] / 0.7, "div_env": lambda s, m: s["x"] / m["env"]["scale"],
    "rdiv": lambda s, m: 3.0 / s["x"], "rdiv_i": lambda s, m: 3 / s["w"],
    "div_t": lambda s, m: s["x"] / m["env"]["delta"],
    "div_xx": lambda s, m: s["x"] / (s["x"] + 1.0), "int_div3": lambda s, m: s["w"] / 3,
    "xsq7": lambda s, m: (s["x"] ** 2) / 7.0,
    "pow07": lambda s, m: abs(s["x"]) ** 0.7, "pow_half": lambda s, m: s["x"] ** 0.5,
    "pow_m1": lambda s, m: s["x"] ** -1,
    "pow_m2": lambda s, m: s["x"] ** -2.0, "pow_25": lambda s, m: s["x"] ** 2.5,
    "pow_n": lambda s, m: s["x"] ** -1.7, "pow_w": lambda s, m: s["w"] ** 0.5,
    "pow_3f": lambda s, m: s["x"] ** 3.0,
    "fdiv_s": lambda s, m: s["x"] // 0.7, "fdiv_sn": lambda s, m: s["x"] // -1.3,
    "fdiv_xx": lambda s, m: s["x"] // (abs(s["x"]) * 0.3 + 0.1),
    "fdiv_env": lambda s, m: s["x"] // m["env"]["scale"],
    "mod_env": lambda s, m: s["x"] % m["env"]["scale"],
    "w_fdiv3": lambda s, m: s["w"] // 3, "w_mod4": lambda s, m: s["w"] % 4,
    "wk_fdiv_m2": lambda s, m: (s["w"] * m["env"]["k"]) // -2,
    "x_mod": lambda s, m: s["x"] % 1.3, "x_modn": lambda s, m: s["x"] % -0.9,
    "log": "log", "min_nan": "min_nan", "max_sz": "max_sz", "min_xy": "min_xy", "clip": "clip",
    "clip_nan": "clip_nan", "exp4": "exp4", "log1": "log1", "sqrt": "sqrt", "sin": "sin",
    "cos": "cos", "tanh": "tanh", "floor": "floor", "ceil": "ceil", "round": "round",
    "sign": "sign", "cast": "cast",
    "where": lambda s, m: (s["b"] * 0 + s["w"] + m["time_step"]),
    "flag": lambda s, m: (s["b"] & (s["x"] > m["env"]["thr"])) | ((s["w"] < 0) ^ m["env"]["on"]),
    "delta": lambda s, m: s["x"] + m["env"]["delta"], "invert": lambda s, m: ~s["b"],
    "half": lambda s, m: s["x"] * 0.5, "sub_t": lambda s, m: s["x"] - s["x"] * 0.5,
    "pow_mhalf": lambda s, m: s["x"] ** -0.5, "pow_w_mhalf": lambda s, m: s["w"] ** -0.5,
    "rsqrt": "rsqrt", "atan2": "atan2", "atan2_w": "atan2_w", "hypot": "hypot",
}


def _jnp_op(name):
    import torch

    from jaxabm_amd import numpy as jnp

    return {
        "rsqrt": lambda s, m: torch.rsqrt(s["x"]),
        "atan2": lambda s, m: jnp.arctan2(s["x"], s["x"] * 0.5 - 1.0),
        "atan2_w": lambda s, m: jnp.arctan2(s["w"], s["x"]),
        "hypot": lambda s, m: jnp.hypot(s["x"], s["x"] * 0.25 + 1.0),
        "log": lambda s, m: jnp.log(s["x"]),
        "min_nan": lambda s, m: jnp.minimum(jnp.log(s["x"]), s["x"] * 0.0),
        "max_sz": lambda s, m: jnp.maximum(s["x"] * 0.0, -(s["x"] * 0.0)),
        "min_xy": lambda s, m: jnp.minimum(s["x"], -s["x"]),
        "clip": lambda s, m: jnp.clip(s["x"], -1.0, 1.5),
        "clip_nan": lambda s, m: jnp.clip(jnp.log(s["x"]), 0.0, 1.0),
        "exp4": lambda s, m: jnp.exp(s["x"] / 4), "log1": lambda s, m: jnp.log(jnp.abs(s["x"]) + 1.0),
        "sqrt": lambda s, m: jnp.sqrt(jnp.abs(s["x"])), "sin": lambda s, m: jnp.sin(s["x"]),
        "cos": lambda s, m: jnp.cos(s["x"]), "tanh": lambda s, m: jnp.tanh(s["x"] * 2),
        "floor": lambda s, m: jnp.floor(s["x"]), "ceil": lambda s, m: jnp.ceil(s["x"]),
        "round": lambda s, m: jnp.round(s["x"] * 3), "sign": lambda s, m: jnp.sign(s["x"]),
        "cast": lambda s, m: jnp.asarray(s["x"] * 10, dtype=jnp.int32),
    }[name]


@pytest.fixture(scope="module")
def wide_inputs(J):
    import torch

    g = torch.Generator(device="cuda").manual_seed(1)
    n = 200_000
    x = (torch.rand(n, device="cuda", generator=g) - 0.5) * torch.exp2(
        torch.randint(-20, 21, (n,), device="cuda", generator=g).float())
    x[:7] = torch.tensor([float("nan"), 0.0, -0.0, float("inf"), -float("inf"), 1e30, -1e-30])
    w = torch.randint(-1000, 1000, (n,), device="cuda", dtype=torch.int32, generator=g)
    b = torch.arange(n, device="cuda") % 3 == 0
    return x, w, b


@pytest.mark.parametrize("name", list(OPS))
def test_single_operation_fused_equals_vmapped(J, wide_inputs, name):
    """One operation on 200,000 values over +-2^20 with NaN, signed zeros,
    infinities and negative integer operands."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import random as jr

    f = OPS[name] if callable(OPS[name]) else _jnp_op(OPS[name])
    x, w, b = wide_inputs
    env = {"shift": 0.25, "scale": 1.5, "k": 3, "thr": 0.1, "on": True,
           "delta": torch.tensor(0.125, device="cuda")}

    class One:
        def init_state(self, cfg, key):
            return {"x": 0.0}

        def update(self, s, ms, cfg, key):
            return {"o": f(s, ms)}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(One(), x.shape[0])
            c.model_config = J.ModelConfig()
            c._states = {"x": x.clone(), "w": w.clone(), "b": b.clone()}
            c.update({"time_step": 2, "env": env}, jr.PRNGKey(5), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0)
            res.append(c.states["o"].clone())
        finally:
            fuse.enable(True)
    a, g = res
    assert a.dtype == g.dtype
    np.testing.assert_array_equal(bits(a), bits(g))


def test_facade_walker_model_fused_equals_vmapped(J):
    """The AgentPy-style model of the reference's basic example
    (examples/basic_example.py:20-140): Agent.setup/step with 2-vectors and an
    env bounds vector, run through the facade with the fused path on and
    off; same positions, colours and records."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp

    class RandomWalker(J.Agent):
        def setup(self):
            return {"position": jnp.array([0.5, 0.5]), "velocity": jnp.array([0.013, 0.007]),
                    "color": 0, "steps_taken": 0}

        def step(self, model_state):
            position, velocity = self._state["position"], self._state["velocity"]
            bounds = model_state.get("env", {}).get("bounds", jnp.array([0.0, 1.0]))
            new_position = position + velocity
            x_bounce = (new_position[0] <= bounds[0]) | (new_position[0] >= bounds[1])
            y_bounce = (new_position[1] <= bounds[0]) | (new_position[1] >= bounds[1])
            new_velocity = velocity * jnp.array([1 - 2 * x_bounce, 1 - 2 * y_bounce])
            new_position = jnp.clip(new_position, bounds[0], bounds[1])
            any_bounce = jnp.logical_or(x_bounce, y_bounce)
            return {"position": new_position, "velocity": new_velocity,
                    "color": jnp.where(any_bounce, 1 - self._state["color"], self._state["color"]),
                    "steps_taken": self._state["steps_taken"] + 1}

    class RandomWalkModel(J.Model):
        def setup(self):
            self.walkers = self.add_agents(self.p.get("n_agents", 50), RandomWalker)
            self.env.add_state("bounds", jnp.array([0.0, 1.0]))

        def step(self):
            self.record("mean_x", float(jnp.mean(self.walkers.position[:, 0])))

    out = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            m = RandomWalkModel({"steps": 120, "n_agents": 64, "seed": 3})
            res = m.run()
            launched = fuse.stats()["launches"] - before
            st = m.walkers.states
            out.append((res["mean_x"], {k: v.clone() for k, v in st.items()}, launched))
        finally:
            fuse.enable(True)
    (ra, sa, la), (rb, sb, lb) = out
    assert la == 0 and lb == 121, (la, lb, fuse.refused())
    assert ra == rb
    for k in sa:
        assert torch.equal(sa[k], sb[k]), k
    assert int(sb["steps_taken"][0]) == 120 and bool((sb["color"] == 1).any())


@pytest.mark.parametrize("k", list(range(1, 17)))
def test_vector_reductions_fused_equal_vmapped(J, k):
    """jnp.sum / jnp.mean / jnp.prod / jnp.linalg.norm / jnp.var / jnp.std
    over a per-agent vector of k float32 values (the orders torch reduces
    them in), integer
    counts and products, on values over +-2^10 with signed zeros (products
    of 16 such values overflow to inf; the scaled product stays finite);
    max / min (whole vector and over its axis) and argmax / argmin of
    vectors full of ties, signed zeros, infinities and NaN, and of int32
    vectors; a power of -0.5."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    g = torch.Generator(device="cuda").manual_seed(k)
    n = 50_000
    p = torch.randn(n, k, device="cuda", generator=g) * torch.exp2(
        torch.randint(-10, 11, (n, k), device="cuda", generator=g).float())
    p[:4, 0] = torch.tensor([-0.0, 0.0, -0.0, 1.0])
    p[:4, 1:] = -0.0
    # extremes: many ties, signed zeros, infinities and NaN
    sel = torch.rand(n, k, device="cuda", generator=g)
    q = torch.randint(-3, 4, (n, k), device="cuda", generator=g).float()
    q[sel < 0.2] = 0.0
    q[(sel >= 0.2) & (sel < 0.4)] = -0.0
    q[(sel >= 0.4) & (sel < 0.43)] = float("inf")
    q[(sel >= 0.43) & (sel < 0.46)] = -float("inf")
    q[(sel >= 0.46) & (sel < 0.48)] = float("nan")
    iq = torch.randint(-5, 6, (n, k), device="cuda", generator=g, dtype=torch.int32)

    class Red:
        def init_state(self, cfg, key):
            return {"q": 0.0}

        def update(self, s, ms, cfg, key):
            v = s["p"]
            return {"s": jnp.sum(v), "s2": jnp.sum(v * v), "m": jnp.mean(v),
                    "c": jnp.sum(v > 0.0), "d": v.sum() - jnp.mean(v * 2.0),
                    "pr": jnp.prod(v * 0.01 + 1.0), "pr0": jnp.prod(v), "pv": v.prod(),
                    "nm": jnp.linalg.norm(v),
                    "ip": jnp.prod((v > 0.0).astype(jnp.int32) + 1),
                    "mx": jnp.max(s["q"]), "mn": jnp.min(s["q"]), "mxm": s["q"].max(),
                    "amx": jnp.max(s["q"], axis=0), "amn": jnp.min(s["q"], axis=0),
                    "am": jnp.argmax(s["q"]), "an": jnp.argmin(s["q"]),
                    "imx": jnp.max(s["iq"]), "iam": jnp.argmax(s["iq"]),
                    "rs": jnp.abs(v) ** -0.5,
                    "vr": jnp.var(v), "vr1": jnp.var(v, ddof=1), "sd": jnp.std(v),
                    "tv": torch.var(v), "ts": torch.std(v, unbiased=False),
                    "vc": torch.var(v, dim=-1, correction=2.5), "qv": jnp.var(s["q"]),
                    "vs": jnp.var(v * 1e-3 + 1000.0)}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(Red(), n)
            c.model_config = J.ModelConfig()
            c._states = {"p": p.clone(), "q": q.clone(), "iq": iq.clone()}
            c.update({"env": {}}, jr.PRNGKey(1), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0), fuse.refused()
            res.append({kk: vv.clone() for kk, vv in c.states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    for kk in a:
        assert a[kk].dtype == b[kk].dtype, kk
        np.testing.assert_array_equal(bits(a[kk]), bits(b[kk]), err_msg=kk)


@pytest.mark.parametrize("k", list(range(1, 16)))
def test_dot_products_fused_equal_vmapped(J, k):
    """jnp.dot, torch.matmul and @ of two per-agent float32 vectors of k <= 15
    values (one multiply-add chain, torch's batched-product order), on
    values over +-2^8 with signed zeros, fused and vmapped alike."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    g = torch.Generator(device="cuda").manual_seed(100 + k)
    n = 30_000
    p = torch.randn(n, k, device="cuda", generator=g) * torch.exp2(
        torch.randint(-8, 9, (n, k), device="cuda", generator=g).float())
    z = torch.rand(n, k, device="cuda", generator=g)
    p[z < 0.1] = 0.0
    p[(z >= 0.1) & (z < 0.2)] = -0.0
    q = torch.randn(n, k, device="cuda", generator=g)

    class Dot:
        def init_state(self, cfg, key):
            return {"x": 0.0}

        def update(self, s, ms, cfg, key):
            v, w = s["p"], s["q"]
            return {"d": jnp.dot(v, w), "m": torch.matmul(w, v), "a": v @ (w * 0.5),
                    "s": jnp.dot(v, v) - jnp.sum(v * v)}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(Dot(), n)
            c.model_config = J.ModelConfig()
            c._states = {"p": p.clone(), "q": q.clone()}
            c.update({"env": {}}, jr.PRNGKey(1), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0), fuse.refused()
            res.append({kk: vv.clone() for kk, vv in c.states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    for kk in a:
        assert a[kk].dtype == b[kk].dtype, kk
        np.testing.assert_array_equal(bits(a[kk]), bits(b[kk]), err_msg=kk)


@pytest.mark.parametrize("k", list(range(1, 17)))
def test_cumulative_sums_fused_equal_vmapped(J, k):
    """jnp.cumsum / torch.cumsum of per-agent float32 vectors (torch's scan
    order) with signed zeros, infinities and NaN, and of int32 and boolean
    vectors (int32 as JAX with x64 off), fused and vmapped alike."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    g = torch.Generator(device="cuda").manual_seed(200 + k)
    n = 30_000
    p = torch.randn(n, k, device="cuda", generator=g) * torch.exp2(
        torch.randint(-10, 11, (n, k), device="cuda", generator=g).float())
    z = torch.rand(n, k, device="cuda", generator=g)
    p[z < 0.1] = 0.0
    p[(z >= 0.1) & (z < 0.25)] = -0.0
    p[(z >= 0.25) & (z < 0.27)] = float("inf")
    p[(z >= 0.27) & (z < 0.28)] = float("nan")
    iq = torch.randint(-2**30, 2**30, (n, k), device="cuda", generator=g, dtype=torch.int32)

    class Scan:
        def init_state(self, cfg, key):
            return {"x": 0.0}

        def update(self, s, ms, cfg, key):
            v = s["p"]
            return {"c": jnp.cumsum(v), "t": torch.cumsum(v * 0.5, dim=0),
                    "i": jnp.cumsum(s["iq"]), "b": jnp.cumsum(v > 0.0),
                    "l": jnp.cumsum(v)[k - 1] - jnp.sum(v)}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(Scan(), n)
            c.model_config = J.ModelConfig()
            c._states = {"p": p.clone(), "iq": iq.clone()}
            c.update({"env": {}}, jr.PRNGKey(1), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0), fuse.refused()
            res.append({kk: vv.clone() for kk, vv in c.states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    assert a["i"].dtype == torch.int32 and a["b"].dtype == torch.int32
    for kk in a:
        assert a[kk].dtype == b[kk].dtype, kk
        np.testing.assert_array_equal(bits(a[kk]), bits(b[kk]), err_msg=kk)


@pytest.mark.parametrize("k", list(range(1, 17)))
def test_sorts_fused_equal_vmapped(J, k):
    """jnp.sort / jnp.argsort / torch.sort(stable=True) of per-agent float32
    vectors full of ties, signed zeros, infinities and NaNs of both signs
    (torch's radix key order), and of int32 and boolean vectors, fused and
    vmapped alike: values bit for bit and indices."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    g = torch.Generator(device="cuda").manual_seed(300 + k)
    n = 30_000
    pool = torch.tensor([-1.0, -0.0, 0.0, 1.0, float("inf"), -float("inf"), 2.0, 0.5],
                        device="cuda")
    p = pool[torch.randint(0, 8, (n, k), device="cuda", generator=g)]
    z = torch.rand(n, k, device="cuda", generator=g)
    neg_nan = torch.tensor([-0x400000], dtype=torch.int32, device="cuda").view(torch.float32)
    p[z < 0.05] = float("nan")
    p[(z >= 0.05) & (z < 0.1)] = neg_nan
    p[z >= 0.9] = torch.randn(int((z >= 0.9).sum()), device="cuda", generator=g)
    iq = torch.randint(-3, 4, (n, k), device="cuda", generator=g, dtype=torch.int32)

    class Sorter:
        def init_state(self, cfg, key):
            return {"x": 0.0}

        def update(self, s, ms, cfg, key):
            v = s["p"]
            return {"s": jnp.sort(v), "a": jnp.argsort(v), "i": jnp.sort(s["iq"]),
                    "ia": jnp.argsort(s["iq"]), "b": jnp.argsort(v > 0.0),
                    "t": torch.sort(v * 2.0, dim=-1, stable=True).indices,
                    "lo": jnp.sort(v)[0] + jnp.sort(v)[k - 1]}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(Sorter(), n)
            c.model_config = J.ModelConfig()
            c._states = {"p": p.clone(), "iq": iq.clone()}
            c.update({"env": {}}, jr.PRNGKey(1), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0), fuse.refused()
            res.append({kk: vv.clone() for kk, vv in c.states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    assert a["a"].dtype == torch.int32 and a["t"].dtype == torch.int64
    for kk in a:
        assert a[kk].dtype == b[kk].dtype, kk
        if kk == "lo":                     # arithmetic: NaN as NaN
            np.testing.assert_array_equal(bits(a[kk]), bits(b[kk]), err_msg=kk)
        elif a[kk].dtype == torch.float32:  # sorted values keep their bits, NaN signs included
            np.testing.assert_array_equal(a[kk].view(torch.int32).cpu().numpy(),
                                          b[kk].view(torch.int32).cpu().numpy(), err_msg=kk)
        else:
            np.testing.assert_array_equal(a[kk].cpu().numpy(), b[kk].cpu().numpy(), err_msg=kk)


def test_vector_logic_concat_flip_roll_take_fused_equal_vmapped(J):
    """jnp.any / all / count_nonzero (NaN counts as true, -0 as zero),
    concatenate, flip, roll and constant takes of per-agent vectors, fused
    and vmapped alike."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    g = torch.Generator(device="cuda").manual_seed(17)
    n = 20_000
    pool = torch.tensor([0.0, -0.0, 1.0, -2.5, float("nan"), float("inf")], device="cuda")
    p = pool[torch.randint(0, 6, (n, 4), device="cuda", generator=g)]
    q = torch.randint(-1, 2, (n, 3), device="cuda", generator=g, dtype=torch.int32)

    class Logic:
        def init_state(self, cfg, key):
            return {"x": 0.0}

        def update(self, s, ms, cfg, key):
            v, w = s["p"], s["q"]
            return {"a": jnp.any(v), "l": jnp.all(v), "c": jnp.count_nonzero(v),
                    "ci": jnp.count_nonzero(w), "ai": jnp.any(w > 0),
                    "cat": jnp.concatenate([v, v * 2.0]), "f": jnp.flip(v),
                    "r": jnp.roll(v, 1), "rn": jnp.roll(w, -1), "t": jnp.take(v, -1),
                    "t0": jnp.take(w, 0)}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            c = J.AgentCollection(Logic(), n)
            c.model_config = J.ModelConfig()
            c._states = {"p": p.clone(), "q": q.clone()}
            c.update({"env": {}}, jr.PRNGKey(1), J.ModelConfig())
            assert fuse.stats()["launches"] - before == (1 if fused else 0), fuse.refused()
            res.append({kk: vv.clone() for kk, vv in c.states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    for kk in a:
        assert a[kk].dtype == b[kk].dtype and a[kk].shape == b[kk].shape, kk
        np.testing.assert_array_equal(bits(a[kk]), bits(b[kk]), err_msg=kk)


def test_vector_at_updates_and_bool_subtraction_fused_equal_vmapped(J):
    """x.at[j].set / add / multiply on per-agent vectors, .astype, and the
    subtraction of a boolean (JAX promotes it; the vmapped path runs it under
    jnp.JaxSemantics), fused and vmapped alike."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    class Updater:
        def init_state(self, cfg, key):
            k1, k2 = jr.split(key)
            return {"p": jr.uniform(k1, (3,), minval=-1.0, maxval=1.0),
                    "q": jr.randint(k2, (2,), 0, 9), "on": jr.bernoulli(k2, 0.5)}

        def update(self, s, ms, cfg, key):
            p, q, on = s["p"], s["q"], s["on"]
            p = p.at[0].set(p[2] * 2.0).at[-1].add(0.25).at[1].multiply(1.0 - on)
            q = q.at[1].add((p[0] > 0).astype(int)).at[0].set(p[1] * 10.0)
            return {"p": p, "q": q, "on": ~on, "f": 1.0 - on, "g": on - 2}

    n = 5000
    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            m = J.JaxModel(config=J.ModelConfig(seed=4))
            m.add_agent_collection("u", J.AgentCollection(Updater(), n))
            m.initialize()
            for _ in range(3):
                m.step()
            assert fuse.stats()["launches"] - before == (4 if fused else 0), fuse.refused()
            res.append({k: v.clone() for k, v in m.agent_collections["u"].states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    assert set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, k
        np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=k)
    assert a["f"].dtype == torch.float32 and a["g"].dtype == torch.int64


def test_gathers_from_agents_arrays_fused_equal_vmapped(J):
    """A voter-style update reading a random other agent's opinion and
    position (model_state['agents_<name>'], jaxabm/model.py:101-116), with
    negative and constant indices and whole-array means and counts, fused
    and vmapped alike; two collections read each other."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp
    from jaxabm_amd import random as jr

    n_v, n_w = 3000, 700

    class Voter:
        def init_state(self, cfg, key):
            k1, k2 = jr.split(key)
            return {"op": jr.uniform(k1), "pos": jr.uniform(k2, (2,)),
                    "tag": jr.randint(k2, (), 0, 4)}

        def update(self, s, ms, cfg, key):
            k1, k2, k3 = jr.split(key, 3)
            a, w = ms["agents_v"], ms["agents_w"]
            j = jr.randint(k1, (), 0, n_v)
            other = a["op"][j]
            near = jnp.abs(other - s["op"]) < ms["env"]["eps"]
            op = jnp.where(near, s["op"] + 0.5 * (other - s["op"]), s["op"])
            mean_op = jnp.mean(a["op"])            # the same for every agent
            pull = jnp.where(jr.uniform(k3) < 0.1, mean_op - op, 0.0)
            return {"op": op + pull, "pos": (s["pos"] + a["pos"][j - n_v]) * 0.5,
                    "tag": a["tag"][j] + w["k"][jr.randint(k2, (), 0, n_w)],
                    "first": a["op"][0] + a["pos"][n_v - 1, -1] + jnp.sum(w["k"] > 4)}

    class Watcher:
        def init_state(self, cfg, key):
            return {"k": jr.randint(key, (), 0, 9), "seen": 0.0}

        def update(self, s, ms, cfg, key):
            return {"k": s["k"], "seen": ms["agents_v"]["op"][jr.randint(key, (), 0, n_v)]}

    res = []
    for fused in (False, True):
        fuse.enable(fused)
        try:
            before = fuse.stats()["launches"]
            m = J.JaxModel(config=J.ModelConfig(seed=9))
            m.add_agent_collection("v", J.AgentCollection(Voter(), n_v))
            m.add_agent_collection("w", J.AgentCollection(Watcher(), n_w))
            m.add_env_state("eps", 0.3)
            m.initialize()
            for _ in range(6):
                m.step()
            assert fuse.stats()["launches"] - before == (14 if fused else 0), fuse.refused()
            res.append({f"{c}.{k}": v.clone() for c in ("v", "w")
                        for k, v in m.agent_collections[c].states.items()})
        finally:
            fuse.enable(True)
    a, b = res
    assert set(a) == set(b)
    for k in a:
        assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, k
        np.testing.assert_array_equal(bits(a[k]), bits(b[k]), err_msg=k)


def test_gather_out_of_range_clamps_as_jax(J):
    """Out-of-range gather indices: negative ones count from the end, the
    rest clamp to the nearest end (JAX's gather).  Run through the fused
    path only (torch's indexing would fault on them)."""
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import random as jr

    n, m = 4000, 37
    src = torch.arange(m, dtype=torch.float32, device="cuda") * 1.5
    key = jr.PRNGKey(12)

    def fn(s, ms, k):
        return {"y": ms["agents_o"]["x"][jr.randint(k, (), -3 * m, 3 * m)]}

    out = fuse.run_agents(fn, n, key, jr.default_mode(), None,
                          {"env": {}, "agents_o": {"x": src}}, fn)
    assert out is not None, fuse.refused()
    j = jr.randint_each(jr.split(key, n), (), -3 * m, 3 * m).cpu().numpy().astype(np.int64)
    jj = np.where(j < 0, j + m, j).clip(0, m - 1)
    np.testing.assert_array_equal(out["y"].cpu().numpy(), src.cpu().numpy()[jj])
    assert (j < -m).any() and (j >= m).any()
