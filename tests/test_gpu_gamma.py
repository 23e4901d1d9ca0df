"""GPU: gamma, log-gamma, beta, lognormal and exponential draws
(jax.random's, called by the reference's examples/models/
advanced_economic_model.py:98,106 and the firms' init_state functions).

* The device sampler against the oracle's restatement on the same keys, bit
  for bit: both evaluate log, log1p, pow and exp with the shared abm_math.h
  functions, so every rejection decision and every value is the same.
* vmapped draws (per-agent keys and parameters) equal single-key draws.
* Agent functions with these draws run as one generated kernel with the
  vmapped program's results, bit for bit.
* ``jnp.nan_to_num`` fused and vmapped."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]


@pytest.fixture(scope="module")
def J():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import jaxabm_amd

    return jaxabm_amd


def _agree(got, want):
    got, want = np.asarray(got, np.float32), np.asarray(want, np.float32)
    assert got.shape == want.shape
    same = (got.view(np.uint32) == want.view(np.uint32)) | (np.isnan(got) & np.isnan(want))
    assert same.all(), (np.count_nonzero(~same), got[~same][:5], want[~same][:5])
    return same


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("a", [0.05, 0.3, 1.0, 2.5, 40.0])
def test_gamma_kernel_matches_oracle(J, mode, a):
    from jaxabm_amd import random as jr

    n = 50_000
    key = jr.PRNGKey(31)
    for log_space in (False, True):
        f = jr.loggamma if log_space else jr.gamma
        got = f(key, a, (n,), mode=mode).cpu().numpy()
        want = O.gamma(O.prng_key(31), a, n, log_space, mode)
        _agree(got, want)


@pytest.mark.parametrize("mode", MODES)
def test_beta_kernel_matches_oracle_with_per_element_parameters(J, mode):
    import torch

    from jaxabm_amd import random as jr

    n = 30_001
    rs = np.random.default_rng(4)
    a = rs.uniform(0.2, 8.0, n).astype(np.float32)
    b = rs.uniform(0.2, 8.0, n).astype(np.float32)
    got = jr.beta(jr.PRNGKey(8), torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda(),
                  mode=mode).cpu().numpy()
    _agree(got, O.beta(O.prng_key(8), a, b, n, mode))
    assert ((got >= 0) & (got <= 1)).all()


def test_edge_cases(J):
    """Empty draws, a = 0 (the boost's power of 1/0 gives 0), NaN and negative
    parameters give NaN, and every lane ends."""
    import torch

    from jaxabm_amd import random as jr

    key = jr.PRNGKey(2)
    assert jr.gamma(key, 1.0, (0,)).shape == (0,)
    a = torch.tensor([0.0, float("nan"), -1.0, 1e-30, 1e30], device="cuda")
    g = jr.gamma(key, a).cpu().numpy()
    want = O.gamma(O.prng_key(2), a.cpu().numpy(), 5)
    _agree(g, want)
    assert np.isnan(g[1]) and np.isnan(g[2]) and g[0] == 0.0


@pytest.mark.parametrize("mode", MODES)
def test_vmapped_draws_equal_single_key_draws(J, mode):
    import torch

    from jaxabm_amd import random as jr

    keys = jr.split(jr.PRNGKey(5), 40, mode=mode)
    a = torch.linspace(0.3, 6.0, 40, device="cuda")

    def one(k, ai):
        return torch.stack([jr.beta(k, ai, 2.0, mode=mode), jr.gamma(k, ai, mode=mode),
                            jr.loggamma(k, 0.5, mode=mode), jr.lognormal(k, 0.5, mode=mode),
                            jr.exponential(k, mode=mode)])

    got = torch.func.vmap(one)(keys, a).cpu().numpy()
    for i in (0, 7, 39):
        want = one(keys[i], a[i]).cpu().numpy()
        np.testing.assert_array_equal(got[i], want)
    vec = torch.func.vmap(lambda k: jr.gamma(k, 2.0, (3,), mode=mode))(keys).cpu().numpy()
    np.testing.assert_array_equal(vec[3], jr.gamma(keys[3], 2.0, (3,), mode=mode).cpu().numpy())


def test_lognormal_and_exponential_closed_forms(J):
    from jaxabm_amd import random as jr

    key = jr.PRNGKey(6)
    z = jr.normal(key, (1000,)).cpu().numpy()
    ln = jr.lognormal(key, 0.5, (1000,)).cpu().numpy()
    # float32 exp and log1p on the device are within an ulp or two of exact
    np.testing.assert_allclose(ln, np.exp((z * np.float32(0.5)).astype(np.float64)), rtol=3e-7)
    u = jr.uniform(key, (1000,)).cpu().numpy()
    ex = jr.exponential(key, (1000,)).cpu().numpy()
    np.testing.assert_allclose(ex, -np.log1p(-u.astype(np.float64)), rtol=3e-7, atol=0)


class Household:
    """A household in the style of the reference's economic example
    (examples/models/advanced_economic_model.py:84-136 init, 138-320 update):
    lognormal savings, beta consumption preference, an employment flip with
    env-dependent probabilities, taxes, consumption, deposits and utility."""

    def init_state(self, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        k1, k2, k3, k4 = jr.split(key, 4)
        savings = 1000.0 * jr.lognormal(k1, sigma=0.5)
        factor = jnp.maximum(0.3, jr.normal(k2) * 0.2 + 1.0)
        return {"savings": savings, "income": 100.0 * factor,
                "cash": savings * 0.3, "deposits": savings * 0.7,
                "propensity": 0.8 * (jr.beta(k3, a=5, b=2) * 0.4 + 0.6),
                "employed": jr.uniform(k4) < 0.95, "productivity": factor,
                "patience": jr.gamma(k4, 2.0) * 0.1, "utility": 0.0}

    def update(self, s, ms, cfg, key):
        from jaxabm_amd import numpy as jnp
        from jaxabm_amd import random as jr

        env = ms["env"]
        k1, k2 = jr.split(key)
        shock = env.get("market", 1.0) * env.get("pandemic", 1.0)
        r = jr.uniform(k1)
        employed = jnp.where(s["employed"], r > 0.02 / shock, r < 0.1 * shock)
        gross = employed * s["productivity"] * env.get("wage", 1.0) + \
            (1.0 - employed) * s["income"] * 0.5
        taxes = gross * env.get("tax", 0.2) * (1.0 + jnp.where(gross > 200.0,
                                                              0.05 * (gross / 100.0 - 2), 0.0))
        net = gross - taxes
        spend = jnp.minimum(net * s["propensity"], s["cash"] + s["deposits"] * 0.3)
        change = jnp.minimum(net * 0.1 - s["deposits"] * 0.1, s["cash"] - spend)
        cash = s["cash"] - spend - jnp.maximum(0, change) + jnp.maximum(0, -change)
        deposits = s["deposits"] + jnp.maximum(0, change) + s["deposits"] * env.get("rate", 0.01)
        noise = jr.exponential(k2) * s["patience"]
        utility = jnp.log1p(spend / env.get("price", 1.0)) + 0.5 * jnp.log1p(deposits / 100) + noise
        return {"savings": cash + deposits, "income": gross, "cash": cash, "deposits": deposits,
                "propensity": s["propensity"], "employed": employed,
                "productivity": s["productivity"], "patience": s["patience"],
                "utility": jnp.nan_to_num(utility, nan=0.0, posinf=1e6)}


def _run(J, agent, n, steps, mode, fused, env):
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import random as jr

    fuse.enable(fused)
    jr.set_threefry_partitionable(mode == O.MODE_PARTITIONABLE)
    try:
        m = J.JaxModel(config=J.ModelConfig(seed=7))
        m.add_agent_collection("h", J.AgentCollection(agent, n))
        for k, v in env.items():
            m.add_env_state(k, v)
        m.initialize()
        init = {k: v.clone() for k, v in m.agent_collections["h"].states.items()}
        for _ in range(steps):
            m.step()
        torch.cuda.synchronize()
        return init, {k: v.clone() for k, v in m.agent_collections["h"].states.items()}
    finally:
        fuse.enable(True)
        jr.set_threefry_partitionable(True)


def _bits(t):
    import torch

    if t.dtype == torch.float32:
        b = t.view(torch.int32).cpu().numpy().copy()
        b[np.isnan(t.cpu().numpy())] = 0x7FC00000
        return b
    return t.cpu().numpy()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [1, 1000, 20_000])
def test_household_fused_equals_vmapped(J, mode, n):
    from jaxabm_amd import fuse

    env = {"market": 0.9, "pandemic": 1.0, "wage": 1.1, "tax": 0.25, "rate": 0.01, "price": 1.02}
    before = fuse.stats()["launches"]
    i_ref, ref = _run(J, Household(), n, 5, mode, False, env)
    mid = fuse.stats()["launches"]
    i_got, got = _run(J, Household(), n, 5, mode, True, env)
    assert mid == before
    assert fuse.stats()["launches"] - mid == 6, fuse.refused()   # init + 5 updates
    for which, (a, b) in {"init": (i_ref, i_got), "end": (ref, got)}.items():
        assert set(a) == set(b)
        for k in a:
            assert a[k].dtype == b[k].dtype and a[k].shape == b[k].shape, (which, k)
            np.testing.assert_array_equal(_bits(a[k]), _bits(b[k]), err_msg=f"{which} {k}")
    p = i_ref["propensity"].cpu().numpy() / 0.8
    assert ((p >= 0.6 - 1e-6) & (p <= 1.0 + 1e-6)).all()
    if n >= 1000:
        assert abs(np.log(i_ref["savings"].cpu().numpy() / 1000.0).std() - 0.5) < 0.05


def test_nan_to_num_fused_equals_vmapped(J):
    import torch

    from jaxabm_amd import fuse
    from jaxabm_amd import numpy as jnp

    class Agent:
        def init_state(self, cfg, key):
            from jaxabm_amd import random as jr

            return {"x": jr.uniform(key, minval=-2.0, maxval=2.0)}

        def update(self, s, ms, cfg, key):
            x = s["x"]
            y = jnp.log(x) / (x - x)           # NaN, +inf and -inf for various x
            return {"x": x, "a": jnp.nan_to_num(y), "b": jnp.nan_to_num(y, nan=0.5, posinf=7.0,
                                                                       neginf=-3.25),
                    "c": jnp.nan_to_num(jnp.log(x)),
                    "i": jnp.nan_to_num(jnp.asarray(x * 3, dtype=jnp.int32))}

    _, ref = _run(J, Agent(), 4000, 1, O.MODE_PARTITIONABLE, False, {})
    before = fuse.stats()["launches"]
    _, got = _run(J, Agent(), 4000, 1, O.MODE_PARTITIONABLE, True, {})
    assert fuse.stats()["launches"] - before == 2, fuse.refused()
    for k in ref:
        assert ref[k].dtype == got[k].dtype
        np.testing.assert_array_equal(_bits(ref[k]), _bits(got[k]), err_msg=k)
    c = ref["c"].cpu().numpy()
    assert np.isfinite(c).all()
    x = torch.tensor([float("nan"), float("inf"), -float("inf"), 1.5], device="cuda")
    np.testing.assert_array_equal(jnp.nan_to_num(x).cpu().numpy(),
                                  np.nan_to_num(x.cpu().numpy()))
