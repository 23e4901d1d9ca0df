"""Utility functions (jaxabm/utils.py:17-176)."""
from __future__ import annotations

from typing import Any, Callable, Dict, List

import numpy as np
import torch


def convert_to_numpy(data: Any) -> Any:
    """Device tensors to numpy arrays, recursively through dicts, lists and
    tuples (jaxabm/utils.py:17-38)."""
    if isinstance(data, dict):
        return {k: convert_to_numpy(v) for k, v in data.items()}
    if isinstance(data, list):
        return [convert_to_numpy(x) for x in data]
    if isinstance(data, tuple):
        return tuple(convert_to_numpy(x) for x in data)
    if isinstance(data, torch.Tensor):
        return data.detach().cpu().numpy()
    return data


def is_valid_params(params: Dict[str, Any], required_keys: List[str]) -> bool:
    return all(k in params for k in required_keys)


def format_time(seconds: float) -> str:
    if seconds < 60:
        return f"{seconds:.2f}s"
    if seconds < 3600:
        minutes = int(seconds / 60)
        return f"{minutes}m {seconds % 60:.2f}s"
    hours = int(seconds / 3600)
    minutes = int((seconds % 3600) / 60)
    return f"{hours}h {minutes}m {seconds % 60:.2f}s"


def mean_over_runs(results_list: List[Dict[str, List[float]]]) -> Dict[str, List[float]]:
    if not results_list:
        return {}
    out = {}
    for key in results_list[0].keys():
        if all(key in r and len(r[key]) == len(results_list[0][key]) for r in results_list):
            arrays = [np.array(convert_to_numpy(r[key])) for r in results_list]
            out[key] = np.mean(arrays, axis=0).tolist()
    return out


def standardize_metrics(metrics: Dict[str, Any]) -> Dict[str, float]:
    out = {}
    for k, v in metrics.items():
        if hasattr(v, "item"):
            out[k] = float(v.item())
        elif isinstance(v, (int, float)):
            out[k] = float(v)
    return out


def run_parallel_simulations(model_factory: Callable[..., Any], param_sets: List[Dict[str, Any]],
                             num_runs: int = 1, seed_offset: int = 0) -> List[Dict[str, Any]]:
    """Every parameter set num_runs times with seed seed_offset + i * num_runs
    + j (jaxabm/utils.py:128-176).  Library models with a server request
    (wealth, network SIR) run as one continuous batch on the batching server
    (``programs.library.run_models``), with the same results."""
    from .core import ModelConfig
    from .programs.library import run_models, servable

    jobs = [(i, j, params, seed_offset + i * num_runs + j)
            for i, params in enumerate(param_sets) for j in range(num_runs)]
    if len(jobs) > 1:
        models = [model_factory(params=p, config=ModelConfig(seed=sd)) for _, _, p, sd in jobs]
        if all(servable(m) for m in models):
            out = []
            for (_, _, params, seed), results in zip(jobs, run_models(models)):
                results["params"] = params
                results["seed"] = seed
                out.append(results)
            return out
    all_results = []
    for i, params in enumerate(param_sets):
        for j in range(num_runs):
            seed = seed_offset + i * num_runs + j
            print(f"Running simulation {i+1}/{len(param_sets)}, run {j+1}/{num_runs}, seed={seed}")
            try:
                results = model_factory(params=params, config=ModelConfig(seed=seed)).run()
                results["params"] = params
                results["seed"] = seed
                all_results.append(results)
            except Exception as e:  # the reference reports and continues
                print(f"Error in simulation {i+1}/{len(param_sets)}, run {j+1}/{num_runs}: {e}")
    return all_results
