"""Model configuration (jaxabm/core.py:55-86) and framework information."""
from __future__ import annotations


class ModelConfig:
    """Configuration for model execution: seed, default step count, history
    tracking and its interval (jaxabm/core.py:55-86)."""

    def __init__(self, seed: int = 0, steps: int = 100, track_history: bool = True,
                 collect_interval: int = 1):
        self.seed = seed
        self.steps = steps
        self.track_history = track_history
        self.collect_interval = collect_interval

    def __repr__(self) -> str:
        return (f"ModelConfig(seed={self.seed}, steps={self.steps}, "
                f"track_history={self.track_history}, collect_interval={self.collect_interval})")


def has_jax() -> bool:
    """This engine does not use JAX: its arrays are GPU tensors and its
    kernels are HIP."""
    return False


def show_info() -> None:
    from . import __version__
    import torch

    print(f"jaxabm_amd v{__version__}: agent-based models on MI355X (HIP kernels)")
    print(f"torch {torch.__version__}, GPU available: {torch.cuda.is_available()}")
    print("Components: Agent, AgentList, Environment, Grid, Network, Model, Results,")
    print("  AgentType, AgentCollection, JaxModel, ModelConfig, SensitivityAnalysis,")
    print("  ModelCalibrator, serving.SimulationServer, programs (SIR, wealth, Schelling,")
    print("  Deffuant)")
