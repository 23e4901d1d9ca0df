"""Run model code written against the reference unchanged.

Reference models import ``jax``, ``jax.numpy``, ``jax.random`` and
``jaxabm`` (jaxabm/agentpy.py:20-23, examples/basic_example.py,
tests/integration/test_integration.py).  JAX is not part of this engine, so
:func:`install` registers those module names for this package:

* ``jaxabm`` -> :mod:`jaxabm_amd`, with ``jaxabm.core``, ``.agent``,
  ``.model``, ``.analysis``, ``.utils`` and ``.agentpy`` / ``.api`` (the
  reference's two names for the AgentPy-style layer);
* ``jax.numpy`` -> :mod:`jaxabm_amd.numpy`, ``jax.random`` ->
  :mod:`jaxabm_amd.random`;
* ``jax`` -> a module with ``numpy``, ``random``, ``vmap``
  (``torch.func.vmap``), ``grad`` (``torch.func.grad``), ``jit`` (returns
  the function: steps are captured as HIP graphs by ``Model.run`` instead),
  ``devices()``, ``__version__`` and ``Array`` (``torch.Tensor``: the type
  the reference's signatures annotate keys and states with, e.g.
  ``key: jax.Array`` in tests/integration/test_integration.py:33).

Call it once, before the model's imports::

    import jaxabm_amd.compat
    jaxabm_amd.compat.install()
    import jax.numpy as jnp          # jaxabm_amd.numpy
    import jaxabm as jx              # jaxabm_amd

It refuses to shadow an importable JAX unless ``force=True``.
"""
from __future__ import annotations

import importlib
import importlib.util
import sys
import types

_JAXABM_SUBMODULES = {"core": "core", "agent": "agent", "model": "model",
                      "analysis": "analysis", "utils": "utils", "agentpy": "agentpy",
                      "api": "agentpy"}


def _jit(fun=None, **_options):
    """jax.jit: the function itself (usable as @jit, @jit(...) or jit(f, ...))."""
    if fun is None:
        return lambda f: f
    return fun


def _devices(backend=None):
    import torch

    if backend in (None, "gpu", "rocm", "cuda") and torch.cuda.is_available():
        return [torch.device("cuda", i) for i in range(torch.cuda.device_count())]
    return [torch.device("cpu")]


def _jax_module() -> types.ModuleType:
    import torch

    from . import __version__
    from . import numpy as jnp
    from . import random as jr

    jax = types.ModuleType("jax", "jax names for jaxabm_amd (jaxabm_amd.compat)")
    jax.numpy = jnp
    jax.random = jr
    jax.vmap = torch.func.vmap
    jax.grad = torch.func.grad
    jax.jit = _jit
    jax.devices = _devices
    jax.Array = torch.Tensor
    jax.__version__ = f"jaxabm_amd-{__version__}"
    jax.__path__ = []  # a package, so `import jax.numpy` resolves through sys.modules
    return jax


_installed: list[str] = []


def install(force: bool = False) -> None:
    """Register ``jax``, ``jax.numpy``, ``jax.random`` and ``jaxabm[.*]`` as
    names of this package's modules (idempotent)."""
    if _installed:
        return
    if not force and ("jax" in sys.modules or importlib.util.find_spec("jax") is not None):
        raise RuntimeError("JAX is importable here: jaxabm_amd.compat.install() would shadow it "
                           "(pass force=True to do so anyway)")
    pkg = importlib.import_module("jaxabm_amd")
    jax = _jax_module()
    names = {"jax": jax, "jax.numpy": jax.numpy, "jax.random": jax.random, "jaxabm": pkg}
    for name, target in _JAXABM_SUBMODULES.items():
        names[f"jaxabm.{name}"] = importlib.import_module(f"jaxabm_amd.{target}")
    for name, mod in names.items():
        sys.modules[name] = mod
        _installed.append(name)
    # `import jaxabm.api` binds the package and reads the attribute
    for name in _JAXABM_SUBMODULES:
        if not hasattr(pkg, name):
            setattr(pkg, name, names[f"jaxabm.{name}"])
            _attributes.append(name)


_attributes: list[str] = []


def uninstall() -> None:
    """Remove the names :func:`install` registered."""
    pkg = sys.modules.get("jaxabm_amd")
    while _attributes:
        name = _attributes.pop()
        if pkg is not None and hasattr(pkg, name):
            delattr(pkg, name)
    while _installed:
        sys.modules.pop(_installed.pop(), None)
