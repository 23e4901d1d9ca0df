"""Host side of the fused agent functions (jaxabm_amd.fuse): tracing, the
dtypes of the generated outputs (torch's promotion with per-agent values as
1-D tensors and env scalars as Python scalars), what is refused, and that
the generated source compiles with hiprtc for gfx950 (no GPU needed)."""
import numpy as np
import pytest
import torch

from jaxabm_amd import fuse
from jaxabm_amd import numpy as jnp
from jaxabm_amd import random as jr

N = 300


def _states():
    return {"x": torch.rand(N), "w": torch.randint(-5, 6, (N,), dtype=torch.int32),
            "alive": torch.rand(N) < 0.5}


def _trace(fn, states=None, ms=None, mode=1):
    key = torch.zeros(2, dtype=torch.int32)
    return fuse.trace(N, key, mode, states, ms, fn, device="cpu")


def test_update_traces_with_torch_dtypes_and_compiles():
    ms = {"time_step": 4, "env": {"rate": 0.5, "n_max": 3, "on": True, "name": "world"}}

    def upd(s, m, k):
        k1, k2 = jr.split(k)
        x = jnp.mod(s["x"] + jr.normal(k1) * 0.01, 1.0)
        w = jnp.where((jr.uniform(k2) < 0.5) & s["alive"], s["w"] - 1, s["w"] + 1)
        return {
            "x": x,
            "w": w + m["time_step"] % 2,           # int32 + Python int
            "f": s["w"] * m["env"]["rate"],        # int32 * Python float -> float32
            "g": s["w"] / 2,                       # true division -> float32
            "h": s["w"] // m["env"]["n_max"],      # floor division stays int32
            "b": s["alive"] | (x > 0.5),
            "bi": s["alive"] + 1,                  # bool + Python int -> int64
            "c": 7, "t": m["time_step"] * 2.5,    # constants: int32, float32
            "r": jr.randint(k1, (), -3, 9) + jr.bernoulli(k2, 0.25),
            "u": jr.uniform(k2, minval=-1.0, maxval=2.0),
            "e": jnp.exp(x) + jnp.sqrt(x) - jnp.tanh(x),
            "cl": jnp.clip(x, 0.2, 0.7), "mx": jnp.maximum(x, s["x"]),
            "cast": jnp.asarray(x * 10, dtype=jnp.int32),
        }

    name, src, tr, outs = _trace(upd, _states(), ms)
    want = {"x": torch.float32, "w": torch.int32, "f": torch.float32, "g": torch.float32,
            "h": torch.int32, "b": torch.bool, "bi": torch.int64, "c": torch.int32,
            "t": torch.float32, "r": torch.int32, "u": torch.float32, "e": torch.float32,
            "cl": torch.float32, "mx": torch.float32, "cast": torch.int32}
    assert {k: v.dtype for k, v in outs.items()} == want
    assert all(v.shape == (N,) for v in outs.values())
    # env Python scalars and the time step are launch parameters, not literals
    assert "world" not in src and "0x1.0000000000000p-1" not in src.split("rate")[0]
    fuse.compile_source(src)


def test_init_traces_and_compiles():
    def init(s, m, k):
        k1, k2 = jr.split(k)
        return {"x": jr.uniform(k1), "w": jr.randint(k2, (), 0, 6), "flag": True}

    name, src, tr, outs = _trace(init)
    assert {k: v.dtype for k, v in outs.items()} == {"x": torch.float32, "w": torch.int32,
                                                      "flag": torch.bool}
    fuse.compile_source(src)


def test_same_function_same_source_for_other_env_values():
    def upd(s, m, k):
        return {"x": s["x"] * m["env"]["rate"] + m["time_step"]}

    a = _trace(upd, _states(), {"time_step": 1, "env": {"rate": 0.5}})[1]
    b = _trace(upd, _states(), {"time_step": 9, "env": {"rate": 0.25}})[1]
    assert a == b


@pytest.mark.parametrize("fn,why", [
    (lambda s, m, k: {"x": s["x"] if s["x"] > 0.5 else -s["x"]}, "control flow"),
    (lambda s, m, k: {"x": jnp.sum(s["x"])}, "sum"),
    (lambda s, m, k: {"x": s["x"] + m["agents_a"]["x"]}, "operand"),
    (lambda s, m, k: {"x": jr.uniform(k, (3, 2))}, "two-dimensional draw"),
    (lambda s, m, k: {"x": float(s["x"])}, "float()"),
    (lambda s, m, k: {"x": s["w"] ** 4}, "integer power"),
    (lambda s, m, k: [s["x"]], "dict"),
])
def test_refused(fn, why):
    ms = {"time_step": 0, "env": {}, "agents_a": {"x": torch.rand(N)}}
    with pytest.raises(fuse.Unsupported):
        _trace(fn, _states(), ms)


def test_vector_limits_refused():
    with pytest.raises(fuse.Unsupported):      # longer than 16
        _trace(lambda s, m, k: {"p": s["p"] + 1.0}, {"p": torch.rand(N, 17)}, {"env": {}})
    with pytest.raises(fuse.Unsupported):      # a norm other than the 2-norm
        _trace(lambda s, m, k: {"r": jnp.linalg.norm(s["p"], ord=1)}, {"p": torch.rand(N, 8)},
               {"env": {}})
    name, src, tr, outs = _trace(lambda s, m, k: {"r": jnp.sum(s["p"] * s["p"]),
                                                  "m": jnp.mean(s["p"]),
                                                  "c": jnp.sum(s["p"] > 0.5)},
                                 {"p": torch.rand(N, 3)}, {"env": {}})
    assert {k: v.dtype for k, v in outs.items()} == {"r": torch.float32, "m": torch.float32,
                                                     "c": torch.int32}   # JAX: int32 sums
    with pytest.raises(fuse.Unsupported):      # vectors of different lengths
        _trace(lambda s, m, k: {"r": s["p"] + s["q"]},
               {"p": torch.rand(N, 2), "q": torch.rand(N, 3)}, {"env": {}})


@pytest.mark.parametrize("k", [1, 8, 11, 16])
def test_vector_sums_products_norms_trace_and_compile(k):
    """Reductions over per-agent vectors of 1..16 values: float32 sums,
    means, products, 2-norms, variances, max / min / argmax / argmin, int32 products
    (JAX x64 off) and a -0.5 power, all in one generated kernel."""
    st = {"p": torch.rand(N, k) - 0.5, "q": torch.randint(-3, 4, (N, k), dtype=torch.int32)}

    def upd(s, m, key):
        v = s["p"]
        return {"s": jnp.sum(v), "m": jnp.mean(v * 2.0), "pr": jnp.prod(v + 1.0),
                "pv": v.prod(), "nm": jnp.linalg.norm(v), "ip": jnp.prod(s["q"]),
                "is": jnp.sum(s["q"]), "mx": jnp.max(v), "amn": jnp.min(v, axis=0),
                "am": jnp.argmax(v), "iam": jnp.argmin(s["q"]), "imx": s["q"].max(),
                "rs": jnp.abs(v) ** -0.5, "rq": torch.rsqrt(jnp.abs(v)),
                "vr": jnp.var(v), "sd": jnp.std(v, ddof=1), "tv": torch.var(v, dim=0)}

    name, src, tr, outs = _trace(upd, st, {"env": {}})
    assert {kk: vv.dtype for kk, vv in outs.items()} == {
        "s": torch.float32, "m": torch.float32, "pr": torch.float32, "pv": torch.float32,
        "nm": torch.float32, "ip": torch.int32, "is": torch.int32, "mx": torch.float32,
        "amn": torch.float32, "am": torch.int32, "iam": torch.int32, "imx": torch.int32,
        "rs": torch.float32, "rq": torch.float32, "vr": torch.float32, "sd": torch.float32,
        "tv": torch.float32}
    assert "abm_wf_merge" in src if k > 1 else "abm_wf_first" in src
    fuse.compile_source(src)


def test_dot_products_trace_compile_and_refuse():
    """jnp.dot / torch.matmul / @ of two per-agent float32 vectors of up to
    15 values is one multiply-add chain; 16 values and a product with an env
    vector are refused (torch orders those otherwise)."""
    st = {"p": torch.rand(N, 15), "q": torch.rand(N, 16)}

    def upd(s, m, key):
        v = s["p"]
        return {"d": jnp.dot(v, v * 2.0), "m": v @ (v + 1.0), "t": torch.dot(v, v)}

    name, src, tr, outs = _trace(upd, st, {"env": {}})
    assert {kk: vv.dtype for kk, vv in outs.items()} == {"d": torch.float32, "m": torch.float32,
                                                         "t": torch.float32}
    assert src.count("__builtin_fmaf(") >= 45
    fuse.compile_source(src)
    with pytest.raises(fuse.Unsupported, match="dot product"):
        _trace(lambda s, m, k: {"d": jnp.dot(s["q"], s["q"])}, st, {"env": {}})
    with pytest.raises(fuse.Unsupported, match="dot product"):
        _trace(lambda s, m, k: {"d": jnp.dot(s["p"], m["env"]["w"])}, st,
               {"env": {"w": torch.rand(15)}}, )


def test_cumulative_sums_trace_and_compile():
    """jnp.cumsum of float32, int32 and boolean per-agent vectors: a
    vector of the input's length, float32 / int32 / int32 (JAX x64 off)."""
    st = {"p": torch.rand(N, 13), "q": torch.randint(-3, 4, (N, 13), dtype=torch.int32)}

    def upd(s, m, key):
        return {"c": jnp.cumsum(s["p"]), "i": jnp.cumsum(s["q"]), "b": jnp.cumsum(s["p"] > 0.5),
                "t": torch.cumsum(s["p"], dim=-1)}

    name, src, tr, outs = _trace(upd, st, {"env": {}})
    assert {kk: (vv.dtype, tuple(vv.shape)) for kk, vv in outs.items()} == {
        "c": (torch.float32, (N, 13)), "i": (torch.int32, (N, 13)), "b": (torch.int32, (N, 13)),
        "t": (torch.float32, (N, 13))}
    fuse.compile_source(src)


def test_sorts_trace_and_compile():
    """jnp.sort / jnp.argsort (stable) of float32 and int32 per-agent
    vectors: values of the input's dtype, argsort int32 (JAX x64 off),
    torch.sort's indices int64; an unstable or descending sort is refused."""
    st = {"p": torch.rand(N, 16), "q": torch.randint(-3, 4, (N, 5), dtype=torch.int32)}

    def upd(s, m, key):
        return {"s": jnp.sort(s["p"]), "a": jnp.argsort(s["p"]), "i": jnp.sort(s["q"]),
                "t": torch.sort(s["q"], stable=True).indices}

    name, src, tr, outs = _trace(upd, st, {"env": {}})
    assert {kk: (vv.dtype, tuple(vv.shape)) for kk, vv in outs.items()} == {
        "s": (torch.float32, (N, 16)), "a": (torch.int32, (N, 16)), "i": (torch.int32, (N, 5)),
        "t": (torch.int64, (N, 5))}
    assert "abm_sort_key" in src
    fuse.compile_source(src)
    for bad in (lambda s, m, k: {"s": torch.sort(s["p"]).values},
                lambda s, m, k: {"s": torch.sort(s["p"], stable=True, descending=True).values}):
        with pytest.raises(fuse.Unsupported):
            _trace(bad, st, {"env": {}})


def test_vector_logic_concat_flip_roll_take_atan2_trace_and_compile():
    """jnp.any / all / count_nonzero, concatenate, flip, roll and a constant
    take of per-agent vectors, arctan2 and hypot: one generated kernel;
    an out-of-range constant take (JAX's fill value) is refused."""
    st = {"p": torch.rand(N, 5) - 0.5, "q": torch.randint(-3, 4, (N, 3), dtype=torch.int32),
          "x": torch.rand(N)}

    def upd(s, m, key):
        v = s["p"]
        return {"a": jnp.any(v > 0.25), "l": jnp.all(v > -0.4), "c": jnp.count_nonzero(s["q"]),
                "cat": jnp.concatenate([v, v * 2.0]), "f": jnp.flip(v), "r": jnp.roll(v, 2),
                "t": jnp.take(v, -1), "y": jnp.arctan2(s["x"], v[0]),
                "h": jnp.hypot(s["x"], 3.0)}

    name, src, tr, outs = _trace(upd, st, {"env": {}})
    assert {kk: (vv.dtype, tuple(vv.shape)) for kk, vv in outs.items()} == {
        "a": (torch.bool, (N,)), "l": (torch.bool, (N,)), "c": (torch.int32, (N,)),
        "cat": (torch.float32, (N, 10)), "f": (torch.float32, (N, 5)),
        "r": (torch.float32, (N, 5)), "t": (torch.float32, (N,)), "y": (torch.float32, (N,)),
        "h": (torch.float32, (N,))}
    assert "atan2f(" in src and "hypotf(" in src
    fuse.compile_source(src)
    with pytest.raises(fuse.Unsupported, match="out of range"):
        _trace(lambda s, m, k: {"t": jnp.take(s["p"], 7)}, st, {"env": {}})


def test_numpy_flip_roll_median_hypot_follow_numpy():
    """jnp.flip / roll / median / hypot on host tensors equal numpy's."""
    rs = np.random.default_rng(2)
    a = rs.standard_normal((4, 6)).astype(np.float32)
    t = torch.from_numpy(a)
    np.testing.assert_array_equal(jnp.flip(t).numpy(), np.flip(a))
    np.testing.assert_array_equal(jnp.flip(t, 1).numpy(), np.flip(a, 1))
    np.testing.assert_array_equal(jnp.roll(t, 3).numpy(), np.roll(a, 3))
    np.testing.assert_array_equal(jnp.roll(t, -2, 1).numpy(), np.roll(a, -2, 1))
    np.testing.assert_allclose(float(jnp.median(t)), np.median(a), rtol=4e-6)
    np.testing.assert_allclose(jnp.median(t, axis=1).numpy(), np.median(a, axis=1), rtol=4e-6)
    assert float(jnp.median(torch.tensor([1, 2, 3, 10], dtype=torch.int32))) == 2.5
    np.testing.assert_allclose(jnp.hypot(t, torch.tensor(2.0)).numpy(),
                               np.hypot(a, np.float32(2.0)), rtol=1e-6)
    assert jnp.hypot(torch.tensor([3]), torch.tensor([4])).item() == 5.0


def test_walker_with_vectors_traces_and_compiles():
    """A bouncing walker in the style of the reference's basic example
    (examples/basic_example.py:20-70): 2-vectors of position and velocity,
    element reads, an env bounds vector, jnp.array of per-agent values and a
    vector draw."""
    st = {"position": torch.rand(N, 2), "velocity": torch.rand(N, 2) * 0.02,
          "color": torch.randint(0, 2, (N,), dtype=torch.int32),
          "steps_taken": torch.zeros(N, dtype=torch.int32)}

    def step(s, m, k):
        position, velocity = s["position"], s["velocity"]
        bounds = m["env"].get("bounds")
        new_position = position + velocity
        x_bounce = (new_position[0] <= bounds[0]) | (new_position[0] >= bounds[1])
        y_bounce = (new_position[1] <= bounds[0]) | (new_position[1] >= bounds[1])
        new_velocity = velocity * jnp.array([1 - 2 * x_bounce, 1 - 2 * y_bounce])
        new_position = jnp.clip(new_position, bounds[0], bounds[1])
        any_bounce = jnp.logical_or(x_bounce, y_bounce)
        return {"position": new_position + jr.normal(k, (2,)) * 0.001,
                "velocity": new_velocity,
                "color": jnp.where(any_bounce, 1 - s["color"], s["color"]),
                "steps_taken": s["steps_taken"] + 1}

    name, src, tr, outs = _trace(step, st, {"env": {"bounds": torch.tensor([0.0, 1.0])},
                                            "time_step": 0})
    assert {k: (v.dtype, tuple(v.shape)) for k, v in outs.items()} == {
        "position": (torch.float32, (N, 2)), "velocity": (torch.float32, (N, 2)),
        "color": (torch.int32, (N,)), "steps_taken": (torch.int32, (N,))}
    fuse.compile_source(src)


def test_integer_sums_are_int32_as_jax_x32():
    """jnp.sum / prod / cumsum of booleans and int32 give int32 with
    wrap-around, as JAX with x64 off (the reference's setting)."""
    x = torch.tensor([2**31 - 1, 1], dtype=torch.int32)
    s = jnp.sum(x)
    assert s.dtype == torch.int32 and int(s) == -2**31
    assert jnp.sum(torch.tensor([True, True, False])).dtype == torch.int32
    assert jnp.prod(torch.tensor([3, 5], dtype=torch.int32)).dtype == torch.int32
    assert jnp.cumsum(torch.tensor([1, 2, 3], dtype=torch.int32)).tolist() == [1, 3, 6]
    assert jnp.sum(torch.tensor([0.5, 0.25])).dtype == torch.float32


def test_gamma_family_and_nan_to_num_trace_and_compile():
    """beta / gamma / loggamma / lognormal / exponential and nan_to_num as
    the reference's economic example uses them
    (examples/models/advanced_economic_model.py:98-106, 1860-1869)."""
    def init(s, m, k):
        k1, k2, k3, k4 = jr.split(k, 4)
        a = jr.lognormal(k1, sigma=0.5) * 1000.0
        return {"a": a, "b": jr.beta(k2, a=5, b=2) * 0.4 + 0.6,
                "g": jr.gamma(k3, 2.5) + jr.loggamma(k3, 0.3),
                "v": jr.beta(k4, 2.0, 5.0, (3,)), "e": jr.exponential(k4),
                "n": jnp.nan_to_num(jnp.log(a - 1000.0), nan=0.1, posinf=5.0)}

    name, src, tr, outs = _trace(init)
    assert {k: (v.dtype, tuple(v.shape)) for k, v in outs.items()} == {
        "a": (torch.float32, (N,)), "b": (torch.float32, (N,)), "g": (torch.float32, (N,)),
        "v": (torch.float32, (N, 3)), "e": (torch.float32, (N,)), "n": (torch.float32, (N,))}
    assert "abm::beta_draw" in src and "abm::gamma_draw" in src
    fuse.compile_source(src)

    def upd(s, m, k):    # a per-agent shape parameter
        return {"x": jr.gamma(k, s["x"] + 0.5), "w": jnp.nan_to_num(s["w"])}

    name, src, tr, outs = _trace(upd, _states(), {"env": {}})
    assert outs["x"].dtype == torch.float32 and outs["w"].dtype == torch.int32
    fuse.compile_source(src)


def test_bool_subtraction_follows_jax():
    """JAX promotes a boolean operand of a subtraction; torch refuses it.
    The traced kernel and the vmapped path (jnp.JaxSemantics) both give
    JAX's values; bool - bool and -bool still raise, as in JAX."""
    from jaxabm_amd.agent import _JAX_SEMANTICS_FNS, vmap_agents

    m = torch.tensor([True, False, True])
    x = torch.tensor([0.5, 1.5, 2.5])

    def fn(a, b):
        return {"y": 1.0 - a, "z": b - a, "w": a - 2}

    out = vmap_agents(fn, m, x, n=3)
    assert out["y"].tolist() == [0.0, 1.0, 0.0] and out["z"].tolist() == [-0.5, 1.5, 1.5]
    assert out["w"].dtype == torch.int64 and out["w"].tolist() == [-1, -2, -1]
    assert fn.__code__ in _JAX_SEMANTICS_FNS
    with jnp.JaxSemantics():
        with pytest.raises(RuntimeError):
            m - m
        with pytest.raises(RuntimeError):
            -m
    name, src, tr, outs = _trace(lambda s, m_, k: {"y": 1.0 - s["alive"], "w": s["alive"] - 2},
                                 _states(), {"env": {}})
    assert outs["y"].dtype == torch.float32 and outs["w"].dtype == torch.int64
    with pytest.raises(fuse.Unsupported):
        _trace(lambda s, m_, k: {"y": -s["alive"]}, _states(), {"env": {}})


def test_astype_reads_dtypes_as_jax_x32():
    """jax.Array.astype on state tensors and traced values: Python int /
    float / bool and 64-bit numpy dtypes read as JAX reads them with x64 off
    (examples/models/advanced_economic_model.py:1493,
    examples/sensitivity/sensitivity_calibration_example.py:137-143)."""
    t = torch.tensor([1.7, -2.5])
    assert t.astype(int).tolist() == [1, -2] and t.astype(int).dtype == torch.int32
    assert t.astype(float).dtype == torch.float32 and t.astype(np.float64).dtype == torch.float32
    assert torch.tensor([True, False]).astype(float).tolist() == [1.0, 0.0]
    assert torch.func.vmap(lambda x: x.astype(int))(t).dtype == torch.int32
    name, src, tr, outs = _trace(lambda s, m, k: {"i": s["x"].astype(int),
                                                  "f": s["alive"].astype(float),
                                                  "v": s["p"].astype(np.int64)},
                                 {**_states(), "p": torch.rand(N, 2)}, {"env": {}})
    assert {k: (v.dtype, tuple(v.shape)) for k, v in outs.items()} == {
        "i": (torch.int32, (N,)), "f": (torch.float32, (N,)), "v": (torch.int32, (N, 2))}


def test_at_updates_follow_jax():
    """x.at[i].set/add/multiply/min/max/divide/get with JAX's meaning: out of
    place, values converted to x's dtype, repeated integer indices
    accumulating (numpy's ufunc.at is the same accumulation)."""
    x = torch.arange(6, dtype=torch.float32)
    i = [1, 1, 3, -1]
    ti = torch.tensor(i)
    ref = np.arange(6, dtype=np.float32)
    for name, ufunc, v in (("add", np.add, 1.5), ("multiply", np.multiply, 2.0),
                           ("min", np.minimum, -1.0), ("max", np.maximum, 10.0)):
        want = ref.copy()
        ufunc.at(want, i, np.float32(v))
        np.testing.assert_array_equal(getattr(x.at[ti], name)(v).numpy(), want, err_msg=name)
    np.testing.assert_array_equal(x.at[ti].divide(2.0).numpy(),
                                  ref / np.array([1, 4, 1, 2, 1, 2], np.float32))
    assert x.at[2].set(9.0).tolist() == [0, 1, 9, 3, 4, 5] and x.tolist() == list(range(6))
    assert x.at[1:3].add(1.0).tolist() == [0, 2, 3, 3, 4, 5]
    assert x.at[[1, 1]].add(1.0).tolist() == [0, 3, 2, 3, 4, 5]
    assert x.at[4].get().item() == 4.0
    assert torch.arange(3, dtype=torch.int32).at[0].add(0.7).tolist() == [0, 1, 2]
    got = torch.func.vmap(lambda v: v.at[0].add(v[1] * 2))(torch.arange(6.).reshape(3, 2))
    assert got.tolist() == [[2, 1], [8, 3], [14, 5]]

    def upd(s, m, k):
        return {"p": s["p"].at[0].set(s["p"][2] * 2).at[-1].add(1.5),
                "q": s["q"].at[1].add(0.7).at[0].multiply(3)}

    name, src, tr, outs = _trace(upd, {"p": torch.rand(N, 3),
                                       "q": torch.randint(0, 5, (N, 2), dtype=torch.int32)},
                                 {"env": {}})
    assert {k: (v.dtype, tuple(v.shape)) for k, v in outs.items()} == {
        "p": (torch.float32, (N, 3)), "q": (torch.int32, (N, 2))}
    fuse.compile_source(src)


def test_gathers_from_agents_arrays_trace():
    """model_state['agents_<name>'][var][j] with a per-agent integer j is a
    gather (jaxabm/model.py:101-116 puts every collection's arrays in
    model_state); whole-array uses are uniform values computed by torch
    before the kernel; an agents array combined with a per-agent value, or
    indexed by a float, is refused."""
    ms = {"env": {}, "time_step": 0,
          "agents_v": {"op": torch.rand(N), "pos": torch.rand(N, 2),
                       "n": torch.randint(0, 3, (N,), dtype=torch.int32)}}

    def upd(s, m, k):
        a = m["agents_v"]
        j = jr.randint(k, (), 0, len(a["op"]))
        return {"x": jnp.where(jr.uniform(k) < 0.5, a["op"][j], s["x"]), "pos": a["pos"][j - N],
                "c": a["pos"][0, -1] + a["n"][j]}

    name, src, tr, outs = _trace(upd, _states(), ms)
    assert {k: (v.dtype, tuple(v.shape)) for k, v in outs.items()} == {
        "x": (torch.float32, (N,)), "pos": (torch.float32, (N, 2)), "c": (torch.float32, (N,))}
    assert "abm_gather_index" in src
    fuse.compile_source(src)
    def uniform(s, m, k):
        a = m["agents_v"]
        return {"x": s["x"] - jnp.mean(a["op"]) + jnp.sum(a["op"] > 0.5) + a["op"].max() +
                (2 * a["op"])[1] + a["pos"][0, 1]}

    name, src, tr, outs = _trace(uniform, _states(), ms)
    assert outs["x"].dtype == torch.float32
    fuse.compile_source(src)
    for bad in (lambda s, m, k: {"x": s["x"] + m["agents_v"]["op"]},
                lambda s, m, k: {"x": m["agents_v"]["op"][s["x"]]}):
        with pytest.raises(fuse.Unsupported):
            _trace(bad, _states(), ms)


def test_jax_gather_indexing_under_vmap():
    """others[j] with a per-agent integer j under vmap (torch would read a
    0-d j on the host, which vmap refuses): vmap_agents reruns the function
    under jnp.JaxSemantics, which gathers with JAX's rule -- negative counts
    from the end, out of range clamps.  Plain tensors follow the same rule
    inside the mode."""
    from jaxabm_amd.agent import _JAX_SEMANTICS_FNS, vmap_agents

    others = torch.arange(10, dtype=torch.float32) * 2
    pos = torch.arange(20, dtype=torch.float32).reshape(10, 2)
    j = torch.tensor([0, 3, -1, -12, 15, 9], dtype=torch.int32)

    def fn(jj):
        return {"o": others[jj], "p": pos[jj], "c": pos[jj, -1], "k": others[jj] + pos[jj, 0]}

    out = vmap_agents(fn, j, n=6)
    jn = np.clip(np.where(j.numpy() < 0, j.numpy() + 10, j.numpy()), 0, 9)
    np.testing.assert_array_equal(out["o"].numpy(), others.numpy()[jn])
    np.testing.assert_array_equal(out["p"].numpy(), pos.numpy()[jn])
    np.testing.assert_array_equal(out["c"].numpy(), pos.numpy()[jn, -1])
    assert fn.__code__ in _JAX_SEMANTICS_FNS
    with jnp.JaxSemantics():
        np.testing.assert_array_equal(others[j].numpy(), others.numpy()[jn])
        assert others[torch.tensor(-3)].item() == 14.0 and others[torch.tensor(40)].item() == 18.0
        assert others[2].item() == 4.0 and others[1:3].tolist() == [2.0, 4.0]
        assert others[others > 15].tolist() == [16.0, 18.0]   # masks are not indices


def test_gradient_run_refuses_host_reads_of_parameters():
    """Inside the calibrator's autograd run a host read of a tensor that
    carries a gradient raises (JAX refuses to concretize under jax.grad), so
    the calibrator differentiates numerically instead of returning a partial
    gradient.  Reads of values without a gradient pass; a branch on a
    comparison follows the taken branch, as torch autograd does."""
    from jaxabm_amd.analysis import _Concrete, _NoConcretization

    p = torch.tensor(0.5, requires_grad=True)
    with _NoConcretization():
        for read in (lambda: float(p), lambda: p.item(), lambda: p.tolist(),
                     lambda: int(p * 2), lambda: float(p * 3.0)):
            with pytest.raises(_Concrete):
                read()
        assert bool(p > 0) is True
        assert float(torch.tensor(2.0)) == 2.0
        assert (p * 2).detach().item() == 1.0


def test_at_updates_and_take_follow_jax_out_of_range_rules():
    """x.at[i] scatters drop integer indices outside x after a negative index
    is counted from the end; x.at[i].get() clamps; jnp.take fills (NaN for
    floats, the most negative value for signed integers) -- JAX's defaults
    (jax.numpy.ndarray.at, jax.numpy.take).  CPU tensors: no GPU fault is
    possible, and the same torch ops run on the GPU."""
    import math

    import torch

    from jaxabm_amd import numpy as jnp

    x = torch.arange(5.0)
    assert x.at[torch.tensor([0, 7, -1, -9])].set(torch.tensor([10.0, 11, 12, 13])).tolist() == \
        [10.0, 1, 2, 3, 12]
    assert x.at[torch.tensor([1, 1, 9])].add(1.0).tolist() == [0.0, 3, 2, 3, 4]
    assert x.at[7].set(3.0).tolist() == x.tolist()
    assert x.at[-1].set(3.0).tolist() == [0.0, 1, 2, 3, 3]
    assert x.at[torch.tensor([1, 1, 9])].multiply(2.0).tolist() == [0.0, 4, 2, 3, 4]
    assert x.at[torch.tensor([2, 8, -6])].max(100.0).tolist() == [0.0, 1, 100, 3, 4]
    assert x.at[torch.tensor([3, 5])].min(-1.0).tolist() == [0.0, 1, 2, -1, 4]
    assert x.at[torch.tensor([4, 6])].power(2).tolist() == [0.0, 1, 2, 3, 16]
    assert x.at[torch.tensor([0, 10])].apply(lambda v: v + 7).tolist() == [7.0, 1, 2, 3, 4]
    assert x.at[torch.tensor([9, -1, -7])].get().tolist() == [4.0, 4, 0]
    m = torch.zeros(3, 4)
    got = m.at[torch.tensor([0, 3, -1]), torch.tensor([1, 1, -2])].add(5.0)
    want = torch.zeros(3, 4)
    want[0, 1] = 5.0
    want[2, 2] = 5.0
    assert torch.equal(got, want)
    # slices and boolean masks keep torch's (in-range) indexing
    assert x.at[1:3].set(9.0).tolist() == [0.0, 9, 9, 3, 4]
    assert x.at[torch.tensor([True, False, True, False, False])].set(1.0).tolist() == \
        [1.0, 1, 1, 3, 4]
    t = jnp.take(x, torch.tensor([0, 9, -1, -7]))
    assert t[0] == 0 and math.isnan(t[1]) and t[2] == 4 and math.isnan(t[3])
    ti = jnp.take(torch.arange(5, dtype=torch.int32), torch.tensor([9, 1]))
    assert ti.tolist() == [torch.iinfo(torch.int32).min, 1]
    t2 = jnp.take(torch.arange(12.0).reshape(3, 4), torch.tensor([0, 5, -1]), axis=1)
    assert t2[:, 0].tolist() == [0.0, 4, 8] and t2[:, 2].tolist() == [3.0, 7, 11]
    assert torch.isnan(t2[:, 1]).all()
    # the input is never modified
    assert x.tolist() == [0.0, 1, 2, 3, 4]
