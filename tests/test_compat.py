"""jaxabm_amd.compat: model code written against the reference's imports
(`import jax`, `import jax.numpy as jnp`, `from jax import random`,
`import jaxabm as jx`, `from jaxabm.<module> import ...`) resolves to this
package.  Run in a child interpreter, so the aliases never leak into this
test process."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CODE = r'''
import jaxabm_amd
import jaxabm_amd.compat as compat
compat.install()
import jax
import jax.numpy as jnp
from jax import random, vmap, jit, grad
import jaxabm as jx
import jaxabm.api
from jaxabm.core import ModelConfig, has_jax, show_info
from jaxabm.model import Model
from jaxabm.agent import AgentType, AgentCollection
from jaxabm.analysis import SensitivityAnalysis, ModelCalibrator
from jaxabm.utils import convert_to_numpy, format_time, run_parallel_simulations
from jaxabm.api import Agent, AgentList, Environment, Grid, Network, Results
from jaxabm.agentpy import Model as AgentPyModel
assert jnp is jaxabm_amd.numpy and random is jaxabm_amd.random and jx is jaxabm_amd
assert jaxabm.api is jaxabm_amd.agentpy and Model is jaxabm_amd.model.Model
assert ModelConfig is jaxabm_amd.ModelConfig and AgentType is jaxabm_amd.AgentType
import torch
@jit
def double(x):
    return x * 2
assert double is not None and jit(double, static_argnums=0) is double
assert torch.equal(vmap(lambda r: r.sum())(torch.ones(3, 4)), torch.full((3,), 4.0))
assert float(grad(lambda x: (x * x).sum())(torch.tensor([3.0]))[0]) == 6.0
assert jax.devices() and jax.__version__.startswith("jaxabm_amd")
# the reference annotates keys and states with jax.Array
# (tests/integration/test_integration.py:33): class bodies must evaluate
assert jax.Array is torch.Tensor
class Walker(AgentType):
    def init_state(self, model_config: ModelConfig, key: jax.Array) -> dict:
        return {"x": random.uniform(key)}
    def update(self, state, model_state, model_config, key: jax.Array) -> dict:
        return state
assert Walker.update.__annotations__["key"] is torch.Tensor
compat.uninstall()
import sys
assert "jax" not in sys.modules and "jaxabm" not in sys.modules
assert not hasattr(jaxabm_amd, "api")
print("ok")
'''


def test_reference_imports_resolve_to_this_package():
    env = dict(os.environ, PYTHONPATH=REPO + os.pathsep + os.environ.get("PYTHONPATH", ""))
    out = subprocess.run([sys.executable, "-c", CODE], cwd=REPO, env=env, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.strip().endswith("ok")


def test_install_refuses_to_shadow_jax():
    code = ("import sys, types; sys.modules['jax'] = types.ModuleType('jax')\n"
            "import jaxabm_amd.compat as c\n"
            "try:\n    c.install()\nexcept RuntimeError as e:\n    print('refused', e)\n")
    env = dict(os.environ, PYTHONPATH=REPO + os.pathsep + os.environ.get("PYTHONPATH", ""))
    out = subprocess.run([sys.executable, "-c", code], cwd=REPO, env=env, capture_output=True,
                         text=True, timeout=300)
    assert out.returncode == 0, out.stderr[-2000:]
    assert "refused" in out.stdout
