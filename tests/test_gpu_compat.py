"""GPU: a model written with the reference's imports (`import jax`,
`import jax.numpy as jnp`, `from jax import random`, `import jaxabm as jx`)
runs unchanged through jaxabm_amd.compat, with the results of the same model
written against jaxabm_amd's own module names."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MODEL = r'''
IMPORTS

class Trader(AgentType):
    def init_state(self, model_config, key: jax.Array):
        k1, k2 = random.split(key)
        return {"cash": random.uniform(k1) * 10.0,
                "risk": random.randint(k2, (), 0, 5)}

    def update(self, state, model_state, model_config, key: jax.Array):
        k1, k2 = random.split(key)
        price = model_state["env"]["price"]
        buy = random.uniform(k1) < 0.5 / (1.0 + state["risk"])
        shock = random.normal(k2) * 0.1
        cash = jnp.where(buy, state["cash"] - price, state["cash"] + price * (1.0 + shock))
        return {"cash": jnp.maximum(cash, 0.0), "risk": state["risk"]}


def update_env(env, agents, params, key):
    new = dict(env)
    new["price"] = env["price"] * (1.0 + 0.01 * (jnp.mean(agents["traders"]["cash"]) - 5.0) / 5.0)
    return new


def metrics(env, agents, params):
    cash = agents["traders"]["cash"]
    return {"mean_cash": jnp.mean(cash), "broke": jnp.sum(cash <= 0.0), "price": env["price"]}


def build_and_run():
    model = Model(params={}, config=ModelConfig(seed=11, steps=40),
                  update_state_fn=update_env, metrics_fn=metrics)
    model.add_env_state("price", 1.0)
    model.add_agent_collection("traders", AgentCollection(Trader(), 5000))
    results = model.run()
    return results, model
'''

REFERENCE_IMPORTS = """
import jax
import jax.numpy as jnp
from jax import random
import jaxabm as jx
from jaxabm.agent import AgentType, AgentCollection
from jaxabm.core import ModelConfig
from jaxabm.model import Model
"""

NATIVE_IMPORTS = """
import types, torch
jax = types.SimpleNamespace(Array=torch.Tensor)   # as compat.install() maps jax.Array
import jaxabm_amd as jx
from jaxabm_amd import numpy as jnp
from jaxabm_amd import random
from jaxabm_amd.agent import AgentType, AgentCollection
from jaxabm_amd.core import ModelConfig
from jaxabm_amd.model import Model
"""


def _run(imports):
    ns = {}
    exec(compile(MODEL.replace("IMPORTS", imports), "<model>", "exec"), ns)
    results, model = ns["build_and_run"]()
    return results, model


def test_reference_style_model_runs_unchanged():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import jaxabm_amd.compat as compat

    native, nm = _run(NATIVE_IMPORTS)
    compat.install()
    try:
        ref, rm = _run(REFERENCE_IMPORTS)
    finally:
        compat.uninstall()
    assert set(ref) == set(native) and len(ref["mean_cash"]) == 40
    for k in ref:
        a = np.array([float(v) for v in ref[k]])
        b = np.array([float(v) for v in native[k]])
        np.testing.assert_array_equal(a, b, err_msg=k)
    sa = rm.agent_collections["traders"].states
    sb = nm.agent_collections["traders"].states
    for k in sa:
        assert torch.equal(sa[k], sb[k]), k
    assert float(ref["broke"][-1]) >= 0 and ref["price"][-1] != ref["price"][0]
