"""JAX-stream PRNG on the GPU: ``jax.random``-shaped functions over the HIP
threefry kernels.

Reference call sites replaced: jaxabm/model.py:46 (PRNGKey), 129/156/164/183
(split), jaxabm/agent.py:115/156 (split(key, N)), jaxabm/agentpy.py:509-513
(randint), jaxabm/analysis.py:81/87 (uniform, permutation).

Keys are int32 tensors of shape (2,) holding the two uint32 words of a JAX
raw key; draws reproduce JAX's bits in either stream mode:

* ``MODE_PARTITIONABLE`` (default) -- ``jax_threefry_partitionable=True``,
  the JAX >= 0.5 default;
* ``MODE_ORIGINAL`` -- the JAX < 0.5 stream.

Batched keys of shape (n, 2) are accepted by the ``*_each`` helpers, which are
the vmap of the scalar draw over the keys (what AgentCollection.update does
with the agent keys, jaxabm/agent.py:156-177).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from ._lib import MODE_ORIGINAL, MODE_PARTITIONABLE, check, lib, ptr, stream_handle

__all__ = [
    "MODE_ORIGINAL", "MODE_PARTITIONABLE", "PRNGKey", "key_data", "split",
    "bits", "uniform", "normal", "bernoulli", "randint", "choice", "permutation", "permutation_each",
    "split_each", "uniform_each", "normal_each", "bernoulli_each",
    "randint_each", "bits_each", "set_threefry_partitionable", "default_mode",
    "gamma", "loggamma", "beta", "lognormal", "exponential",
]

_default_mode = MODE_PARTITIONABLE


def set_threefry_partitionable(flag: bool) -> None:
    """Select the JAX stream (mirrors ``jax.config.update('jax_threefry_partitionable', flag)``)."""
    global _default_mode
    _default_mode = MODE_PARTITIONABLE if flag else MODE_ORIGINAL


def default_mode() -> int:
    return _default_mode


def _mode(mode):
    return _default_mode if mode is None else int(mode)


def _device(device=None):
    _lib.require_gpu()
    return torch.device(device) if device is not None else torch.device("cuda")


def _shape(shape) -> tuple:
    if shape is None:
        return ()
    if isinstance(shape, int):
        return (shape,)
    return tuple(int(s) for s in shape)


def _numel(shape) -> int:
    n = 1
    for s in shape:
        n *= s
    return n


def _as_key(key) -> torch.Tensor:
    if isinstance(key, torch.Tensor):
        if key.shape[-1:] != (2,):
            raise ValueError(f"a PRNG key has trailing dimension 2, got shape {tuple(key.shape)}")
        if not key.is_cuda:
            key = key.to("cuda")
        return key.to(torch.int32).contiguous()
    arr = np.asarray(key, dtype=np.uint32)
    if arr.shape[-1:] != (2,):
        raise ValueError(f"a PRNG key has trailing dimension 2, got shape {arr.shape}")
    return torch.from_numpy(arr.view(np.int32).copy()).to(_device())


def PRNGKey(seed: int, device=None) -> torch.Tensor:
    """``jax.random.PRNGKey(seed)`` for 32-bit seeds: ``[0, seed mod 2^32]``."""
    s = int(seed) & 0xFFFFFFFF
    host = np.array([0, s], dtype=np.uint32).view(np.int32)
    return torch.from_numpy(host.copy()).to(_device(device))


key = PRNGKey


def key_data(key) -> np.ndarray:
    """Host uint32 view of a key (or key array)."""
    if isinstance(key, torch.Tensor):
        return key.detach().to("cpu").contiguous().numpy().view(np.uint32).copy()
    return np.asarray(key, dtype=np.uint32)


def _single(key) -> torch.Tensor:
    k = _as_key(key)
    if k.dim() != 1:
        raise ValueError("expected a single key of shape (2,); use the *_each helpers for key arrays")
    return k


def _split_impl(key, num: int = 2, mode=None) -> torch.Tensor:
    k = _single(key)
    out = torch.empty((int(num), 2), dtype=torch.int32, device=k.device)
    check(lib().abm_split(ptr(k), int(num), ptr(out), _mode(mode), stream_handle()), "split")
    return out


def _bits_impl(key, shape=(), mode=None) -> torch.Tensor:
    k = _single(key)
    shp = _shape(shape)
    out = torch.empty(shp, dtype=torch.int32, device=k.device)
    check(lib().abm_random_bits32(ptr(k), _numel(shp), ptr(out), _mode(mode), stream_handle()), "bits")
    return out


def _uniform_impl(key, shape=(), minval=0.0, maxval=1.0, mode=None) -> torch.Tensor:
    k = _single(key)
    shp = _shape(shape)
    out = torch.empty(shp, dtype=torch.float32, device=k.device)
    check(lib().abm_uniform_f32(ptr(k), _numel(shp), float(minval), float(maxval), ptr(out),
                                _mode(mode), stream_handle()), "uniform")
    return out


def _normal_impl(key, shape=(), mode=None) -> torch.Tensor:
    k = _single(key)
    shp = _shape(shape)
    out = torch.empty(shp, dtype=torch.float32, device=k.device)
    check(lib().abm_normal_f32(ptr(k), _numel(shp), ptr(out), _mode(mode), stream_handle()), "normal")
    return out


def _bernoulli_impl(key, p=0.5, shape=(), mode=None) -> torch.Tensor:
    k = _single(key)
    shp = _shape(shape)
    out = torch.empty(shp, dtype=torch.uint8, device=k.device)
    check(lib().abm_bernoulli_u8(ptr(k), _numel(shp), float(p), ptr(out), _mode(mode),
                                 stream_handle()), "bernoulli")
    return out.bool()


def _randint_impl(key, shape, minval, maxval, mode=None) -> torch.Tensor:
    k = _single(key)
    shp = _shape(shape)
    lo = int(np.clip(int(minval), -2**31, 2**31 - 1))
    hi = int(np.clip(int(maxval), -2**31, 2**31 - 1))
    out = torch.empty(shp, dtype=torch.int32, device=k.device)
    check(lib().abm_randint_i32(ptr(k), _numel(shp), lo, hi, ptr(out), _mode(mode),
                                stream_handle()), "randint")
    return out


# ---------------------------------------------------------------- vmap-able draws
# The draws are torch custom ops with vmap rules: called with one key they run
# the single-key kernels; inside torch.func.vmap (what AgentCollection uses to
# map a per-agent function over the agents, as jax.vmap does in
# jaxabm/agent.py:115-177) the batched key array goes to the per-key kernels
# in one launch -- the same bits JAX produces for a vmapped draw.
def _batched_keys(info, dim, key) -> torch.Tensor:
    if dim is None:
        return key.reshape(1, 2).expand(info.batch_size, 2).contiguous()
    return key.movedim(dim, 0).contiguous()


@torch.library.custom_op("jaxabm::split", mutates_args=(), device_types="cuda")
def _op_split(key: torch.Tensor, num: int, mode: int) -> torch.Tensor:
    return _split_impl(key, num, mode)


@_op_split.register_fake
def _(key, num, mode):
    return key.new_empty((num, 2))


@_op_split.register_vmap
def _(info, in_dims, key, num, mode):
    return split_each(_batched_keys(info, in_dims[0], key), num, mode), 0


_DRAW_DTYPES = {0: torch.int32, 1: torch.float32, 2: torch.float32}


@torch.library.custom_op("jaxabm::draw", mutates_args=(), device_types="cuda")
def _op_draw(key: torch.Tensor, shape: list[int], dist: int, lo: float, hi: float,
             mode: int) -> torch.Tensor:
    if dist == 0:
        return _bits_impl(key, tuple(shape), mode)
    if dist == 1:
        return _uniform_impl(key, tuple(shape), lo, hi, mode)
    return _normal_impl(key, tuple(shape), mode)


@_op_draw.register_fake
def _(key, shape, dist, lo, hi, mode):
    return key.new_empty(tuple(shape), dtype=_DRAW_DTYPES[dist])


@_op_draw.register_vmap
def _(info, in_dims, key, shape, dist, lo, hi, mode):
    k = _batched_keys(info, in_dims[0], key)
    return _draw_each(k, tuple(shape), dist, lo, hi, _DRAW_DTYPES[dist], mode), 0


@torch.library.custom_op("jaxabm::randint", mutates_args=(), device_types="cuda")
def _op_randint(key: torch.Tensor, shape: list[int], lo: int, hi: int, mode: int) -> torch.Tensor:
    return _randint_impl(key, tuple(shape), lo, hi, mode)


@_op_randint.register_fake
def _(key, shape, lo, hi, mode):
    return key.new_empty(tuple(shape), dtype=torch.int32)


@_op_randint.register_vmap
def _(info, in_dims, key, shape, lo, hi, mode):
    return randint_each(_batched_keys(info, in_dims[0], key), tuple(shape), lo, hi, mode), 0


def _gamma_each_impl(keys: torch.Tensor, a: torch.Tensor, b: torch.Tensor, kind: int,
                     mode: int) -> torch.Tensor:
    """keys (n, 2); a, b (n, *shape) float32 -> (n, *shape)."""
    k, a, b = keys.contiguous(), a.contiguous(), b.contiguous()
    out = torch.empty(a.shape, dtype=torch.float32, device=k.device)
    n = k.shape[0]
    per = a.numel() // n if n else 0
    check(lib().abm_gamma_each(ptr(k), n, per, ptr(a), ptr(b), kind, ptr(out), mode,
                               stream_handle()), "gamma")
    return out


@torch.library.custom_op("jaxabm::gamma", mutates_args=(), device_types="cuda")
def _op_gamma(key: torch.Tensor, a: torch.Tensor, b: torch.Tensor, kind: int,
              mode: int) -> torch.Tensor:
    out = _gamma_each_impl(key.reshape(1, 2), a.reshape(1, *a.shape), b.reshape(1, *b.shape),
                           kind, mode)
    return out.reshape(a.shape)


@_op_gamma.register_fake
def _(key, a, b, kind, mode):
    return a.new_empty(a.shape)


@_op_gamma.register_vmap
def _(info, in_dims, key, a, b, kind, mode):
    k = _batched_keys(info, in_dims[0], key)

    def batched(t, d):
        return t.movedim(d, 0) if d is not None else t.unsqueeze(0).expand(info.batch_size, *t.shape)

    return _gamma_each_impl(k, batched(a, in_dims[1]), batched(b, in_dims[2]), kind, mode), 0


def _key_arg(key) -> torch.Tensor:
    if isinstance(key, torch.Tensor):
        if key.shape[-1:] != (2,):
            raise ValueError(f"a PRNG key has trailing dimension 2, got shape {tuple(key.shape)}")
        if key.dtype != torch.int32:
            key = key.to(torch.int32)
        return key if key.is_cuda else key.to(_device())
    return _as_key(key)


# A traced key (jaxabm_amd.fuse: a per-agent function being lowered to one
# kernel) takes the draws below to the trace, with the same constructions.
def _traced_key(key) -> bool:
    return getattr(key, "_abm_symkey", False) is True


def _traced_scalar_shape(key, shape, mode):
    from .fuse import SymKey, Unsupported

    if not isinstance(key, SymKey):
        raise Unsupported("a draw from a key array")
    if len(_shape(shape)) > 1:
        raise Unsupported("a draw of more than one dimension inside a per-agent function")
    if _mode(mode) != key.tr.mode:
        raise Unsupported("a draw in another stream mode")


def split(key, num: int = 2, mode=None) -> torch.Tensor:
    """``jax.random.split(key, num)`` -> (num, 2)."""
    if _traced_key(key):
        _traced_scalar_shape(key, (), mode)
        return key.split(int(num))
    return torch.ops.jaxabm.split(_key_arg(key), int(num), _mode(mode))


def bits(key, shape=(), mode=None) -> torch.Tensor:
    """``jax.random.bits(key, shape, uint32)`` (returned as int32 storage)."""
    if _traced_key(key):
        _traced_scalar_shape(key, shape, mode)
        return key.bits(_shape(shape))
    return torch.ops.jaxabm.draw(_key_arg(key), list(_shape(shape)), 0, 0.0, 0.0, _mode(mode))


def uniform(key, shape=(), dtype=torch.float32, minval=0.0, maxval=1.0, mode=None) -> torch.Tensor:
    """``jax.random.uniform(key, shape, float32, minval, maxval)``.  Tensor
    bounds (e.g. per agent under vmap) scale a [0, 1) draw with torch ops."""
    if dtype not in (torch.float32, None):
        raise NotImplementedError("uniform: float32 only (the reference runs with jax_enable_x64 off)")
    if _traced_key(key):
        _traced_scalar_shape(key, shape, mode)
        return key.uniform(minval, maxval, _shape(shape))
    k = _key_arg(key)
    if isinstance(minval, torch.Tensor) or isinstance(maxval, torch.Tensor):
        f = torch.ops.jaxabm.draw(k, list(_shape(shape)), 1, 0.0, 1.0, _mode(mode))
        def bound(v):
            if isinstance(v, torch.Tensor):
                return v.to(device=f.device, dtype=torch.float32)
            return torch.full((), float(v), dtype=torch.float32, device=f.device)

        lo, hi = bound(minval), bound(maxval)
        return torch.maximum(lo, f * (hi - lo) + lo)
    return torch.ops.jaxabm.draw(k, list(_shape(shape)), 1, float(minval), float(maxval),
                                 _mode(mode))


def normal(key, shape=(), dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.normal(key, shape, float32)``."""
    if _traced_key(key):
        _traced_scalar_shape(key, shape, mode)
        return key.normal(_shape(shape))
    return torch.ops.jaxabm.draw(_key_arg(key), list(_shape(shape)), 2, 0.0, 0.0, _mode(mode))


def bernoulli(key, p=0.5, shape=(), mode=None) -> torch.Tensor:
    """``jax.random.bernoulli(key, p, shape)`` = ``uniform(key, shape) < p``."""
    if isinstance(p, torch.Tensor) or getattr(p, "_abm_sym", False) is True:
        shp = _shape(shape) if shape not in ((), None) else tuple(p.shape)
        return uniform(key, shp, mode=mode) < p.to(torch.float32)
    return uniform(key, shape, mode=mode) < float(p)


def randint(key, shape, minval, maxval, dtype=torch.int32, mode=None) -> torch.Tensor:
    """``jax.random.randint(key, shape, minval, maxval, int32)``."""
    lo = int(np.clip(int(minval), -2**31, 2**31 - 1))
    hi = int(np.clip(int(maxval), -2**31, 2**31 - 1))
    if _traced_key(key):
        _traced_scalar_shape(key, shape, mode)
        return key.randint(lo, hi, _shape(shape))
    return torch.ops.jaxabm.randint(_key_arg(key), list(_shape(shape)), lo, hi, _mode(mode))


def _param_shape(params, shape) -> tuple:
    if shape is not None:
        return _shape(shape)
    return tuple(torch.broadcast_shapes(*(tuple(getattr(v, "shape", ())) for v in params)))


def _param(v, shape: tuple, device) -> torch.Tensor:
    """A draw parameter as float32 (jax.random converts it to the draw's
    dtype) broadcast to the draw's shape."""
    if isinstance(v, torch.Tensor):
        t = v.to(device=device, dtype=torch.float32)
    else:
        t = torch.full((), float(v), dtype=torch.float32, device=device)
    return t.broadcast_to(shape)


def _gamma_family(key, a, b, shape, kind: int, dtype, mode):
    if dtype not in (torch.float32, None):
        raise NotImplementedError("gamma/beta: float32 only (the reference runs with "
                                  "jax_enable_x64 off)")
    params = (a,) if b is None else (a, b)
    if _traced_key(key):
        from .fuse import SymVec

        shp = () if shape is None else _shape(shape)
        if shape is None:
            lens = {len(v) for v in params if isinstance(v, SymVec)}
            shp = (lens.pop(),) if len(lens) == 1 else shp
        _traced_scalar_shape(key, shp, mode)
        return key.gamma(a, b, kind, shp)
    k = _key_arg(key)
    shp = _param_shape(params, shape)
    at = _param(a, shp, k.device)
    bt = _param(b, shp, k.device) if b is not None else at
    return torch.ops.jaxabm.gamma(k, at, bt, kind, _mode(mode))


def gamma(key, a, shape=None, dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.gamma(key, a, shape, float32)``: Marsaglia-Tsang rejection
    sampling with jax.random's key layout (element j of a draw of k elements
    from key ``split(key, k)[j]``); ``shape`` defaults to ``a``'s."""
    return _gamma_family(key, a, None, shape, _lib.GAMMA, dtype, mode)


def loggamma(key, a, shape=None, dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.loggamma(key, a, shape, float32)``: log of a gamma draw,
    computed in log space (accurate for small ``a``)."""
    return _gamma_family(key, a, None, shape, _lib.LOGGAMMA, dtype, mode)


def beta(key, a, b, shape=None, dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.beta(key, a, b, shape, float32)`` (called by the reference's
    examples/models/advanced_economic_model.py:106,351,635,958,962): loggamma
    draws of a and b from the two halves of ``split(key)``, combined as
    ``exp(la - m) / (exp(la - m) + exp(lb - m))``, ``m = max(la, lb)``."""
    return _gamma_family(key, a, b, shape, _lib.BETA, dtype, mode)


def lognormal(key, sigma=1.0, shape=None, dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.lognormal(key, sigma, shape, float32)`` =
    ``exp(normal(key, shape) * sigma)`` (examples/models/
    advanced_economic_model.py:98,343,627,950)."""
    if dtype not in (torch.float32, None):
        raise NotImplementedError("lognormal: float32 only")
    shp = _param_shape((sigma,), shape) if not _traced_key(key) or shape is not None else ()
    z = normal(key, shp, mode=mode)
    if isinstance(sigma, torch.Tensor):
        sigma = sigma.to(torch.float32)
    return torch.exp(z * sigma)


def exponential(key, shape=(), dtype=torch.float32, mode=None) -> torch.Tensor:
    """``jax.random.exponential(key, shape, float32)`` = ``-log1p(-uniform)``."""
    if dtype not in (torch.float32, None):
        raise NotImplementedError("exponential: float32 only")
    u = uniform(key, shape, mode=mode)
    return torch.neg(torch.log1p(torch.neg(u)))


def choice(key, a, shape=(), mode=None):
    """``jax.random.choice(key, a, shape)`` with replacement:
    ``a[randint(key, shape, 0, len(a))]`` (a an int n or a 1-D tensor)."""
    if isinstance(a, int):
        return randint(key, shape, 0, a, mode=mode)
    idx = randint(key, shape, 0, a.shape[0], mode=mode)
    return a[idx.long()]


def permutation(key, x, mode=None) -> torch.Tensor:
    """``jax.random.permutation(key, n)`` for an integer n, or of a 1-D tensor."""
    if _traced_key(key):
        from .fuse import Unsupported

        raise Unsupported("a permutation inside a per-agent function")
    k = _single(key)
    if isinstance(x, torch.Tensor):
        n = x.shape[0]
    else:
        n = int(x)
    out = torch.empty((n,), dtype=torch.int32, device=k.device)
    if n:
        nbytes = int(lib().abm_permutation_workspace(n))
        ws = torch.empty((nbytes,), dtype=torch.uint8, device=k.device)
        check(lib().abm_permutation(ptr(k), n, ptr(out), ptr(ws), nbytes, _mode(mode),
                                    stream_handle()), "permutation")
    if isinstance(x, torch.Tensor):
        return x[out.long()]
    return out


def permutation_each(keys, n: int, mode=None) -> torch.Tensor:
    """``vmap(lambda k: permutation(k, n))(keys)`` for keys (count, 2) ->
    (count, n) int32, computed by one batched segmented sort."""
    k = _as_key(keys).reshape(-1, 2).contiguous()
    count = k.shape[0]
    out = torch.empty((count, n), dtype=torch.int32, device=k.device)
    if n and count:
        nbytes = int(lib().abm_permutation_batch_workspace(n, count))
        ws = torch.empty((nbytes,), dtype=torch.uint8, device=k.device)
        check(lib().abm_permutation_batch(ptr(k), 2, count, n, ptr(out), ptr(ws), nbytes,
                                          _mode(mode), stream_handle()), "permutation_each")
    return out


# ---------------------------------------------------------------- vmapped draws
def _keys(keys) -> torch.Tensor:
    k = _as_key(keys)
    return k.reshape(-1, 2)


def split_each(keys, num: int = 2, mode=None) -> torch.Tensor:
    """``vmap(lambda k: split(k, num))(keys)`` -> (n, num, 2)."""
    k = _keys(keys)
    out = torch.empty((k.shape[0], int(num), 2), dtype=torch.int32, device=k.device)
    check(lib().abm_split_each(ptr(k), k.shape[0], int(num), ptr(out), _mode(mode),
                               stream_handle()), "split_each")
    return out


def _draw_each(keys, shape, dist, lo, hi, dtype, mode):
    k = _keys(keys)
    shp = _shape(shape)
    out = torch.empty((k.shape[0], *shp), dtype=dtype, device=k.device)
    check(lib().abm_draw_each(ptr(k), k.shape[0], _numel(shp), dist, float(lo), float(hi),
                              ptr(out), _mode(mode), stream_handle()), "draw_each")
    return out


def bits_each(keys, shape=(), mode=None):
    return _draw_each(keys, shape, 0, 0.0, 0.0, torch.int32, mode)


def uniform_each(keys, shape=(), minval=0.0, maxval=1.0, mode=None):
    return _draw_each(keys, shape, 1, minval, maxval, torch.float32, mode)


def normal_each(keys, shape=(), mode=None):
    return _draw_each(keys, shape, 2, 0.0, 0.0, torch.float32, mode)


def bernoulli_each(keys, p=0.5, shape=(), mode=None):
    return _draw_each(keys, shape, 3, p, 0.0, torch.uint8, mode).bool()


def randint_each(keys, shape, minval, maxval, mode=None):
    k = _keys(keys)
    shp = _shape(shape)
    out = torch.empty((k.shape[0], *shp), dtype=torch.int32, device=k.device)
    check(lib().abm_randint_each(ptr(k), k.shape[0], _numel(shp), int(minval), int(maxval),
                                 ptr(out), _mode(mode), stream_handle()), "randint_each")
    return out


# ---------------------------------------------------------------- host key math
def host_split(key, num: int = 2, mode=None) -> np.ndarray:
    """Split a host uint32[2] key on the host (no device work) -> (num, 2)."""
    k = np.ascontiguousarray(np.asarray(key, dtype=np.uint32).reshape(2))
    out = np.zeros(2 * int(num), np.uint32)
    check(lib().abm_host_split(k, int(num), out, _mode(mode)), "host_split")
    return out.reshape(int(num), 2)


def host_prng_key(seed: int) -> np.ndarray:
    return np.array([0, int(seed) & 0xFFFFFFFF], dtype=np.uint32)


def step_key_table(rng: np.ndarray, steps: int, n_collections: int, has_env: bool,
                   split_env: bool, mode=None) -> np.ndarray:
    """Host key schedule of ``steps`` JaxModel.step calls; advances ``rng`` in place."""
    width = 2 * n_collections + ((4 if split_env else 2) if has_env else 0)
    out = np.zeros(max(steps, 1) * max(width, 1), np.uint32)
    check(lib().abm_step_keys(rng, int(steps), int(n_collections), int(bool(has_env)),
                              int(bool(split_env)), _mode(mode), out), "step_keys")
    return out[: steps * width].reshape(steps, width)
