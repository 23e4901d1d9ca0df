"""jaxabm_amd: agent-based models on AMD Instinct MI355X.

The API of the reference (a11to1n3/JaxABM) on hand-written HIP kernels:

* AgentPy-style façade: Agent, AgentList, Environment, Grid, Network, Model,
  Results, Parameter, Sample, SensitivityAnalyzer, ModelCalibrator;
* core: AgentType, AgentCollection, JaxModel, ModelConfig;
* analysis: SensitivityAnalysis, ModelCalibrator;
* utilities: convert_to_numpy, format_time, run_parallel_simulations;
* ``jaxabm_amd.random`` (JAX threefry streams) and ``jaxabm_amd.numpy``
  (jnp-shaped array functions) for agent code;
* fused programs (``jaxabm_amd.programs``): network SIR, wealth exchange,
  Schelling, differentiable Deffuant; the batching server
  (``jaxabm_amd.serving``).

Every compute path is the HIP library (``libjaxabm_hip.so``); importing the
package does not touch the GPU.
"""
__version__ = "0.1.5"

from .core import ModelConfig, has_jax, show_info  # noqa: E402
from .agent import AgentCollection, AgentType  # noqa: E402
from .model import JaxModel  # noqa: E402
from .utils import convert_to_numpy, format_time, run_parallel_simulations  # noqa: E402
from .agentpy import (Agent, AgentList, Environment, Grid, Model, ModelCalibrator,  # noqa: E402
                      Network, Parameter, Results, Sample, SensitivityAnalyzer)
from .analysis import SensitivityAnalysis  # noqa: E402


def jax_available() -> bool:
    """The engine's components are always available (it does not need JAX)."""
    return True


__all__ = [
    "Agent", "AgentList", "Environment", "Grid", "Network", "Model", "Results", "Parameter",
    "Sample", "SensitivityAnalyzer", "AgentType", "AgentCollection", "JaxModel", "ModelConfig",
    "SensitivityAnalysis", "ModelCalibrator", "convert_to_numpy", "format_time",
    "run_parallel_simulations", "has_jax", "jax_available", "show_info",
]
