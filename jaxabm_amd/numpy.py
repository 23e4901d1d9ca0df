"""``jax.numpy``-shaped array functions over torch, for agent and model code.

Models written against the reference use ``jax.numpy`` inside ``Agent.step``,
``AgentType.update`` and the env/metrics callbacks (jaxabm/agent.py:19-51,
jaxabm/model.py:29-30).  This namespace gives the same names on GPU tensors,
with JAX's default dtypes (float32, int32: the reference runs with
jax_enable_x64 off), so such code runs with ``import jaxabm_amd.numpy as jnp``.
Every function is a torch op, so per-agent functions run under
``torch.func.vmap`` exactly as under ``jax.vmap``.

JAX behaviours torch lacks are added where torch has no name for them:
``x.at[i].set(v)`` (and ``add``, ``multiply``, ...), ``x.astype(float)``,
and, inside :class:`JaxSemantics`, the promotion of a boolean operand of a
subtraction and JAX's gather indexing by integer tensors.
"""
from __future__ import annotations

import builtins
import math
import warnings

import numpy as _np
import torch

float32 = torch.float32
float64 = torch.float64
int32 = torch.int32
int64 = torch.int64
uint8 = torch.uint8
uint32 = torch.int32  # 32-bit storage; random bits come as int32 bit patterns
bool_ = torch.bool
pi = math.pi
inf = math.inf
nan = math.nan
e = math.e
ndarray = torch.Tensor


def _dev():
    from ._lib import require_gpu

    require_gpu()
    return torch.device("cuda")


def _default_dtype(value):
    if isinstance(value, bool):
        return torch.bool
    if isinstance(value, int):
        return torch.int32
    if isinstance(value, float):
        return torch.float32
    return None


def jax_dtype(dt) -> torch.dtype:
    """A dtype argument read as JAX reads it with jax_enable_x64 off (the
    reference's setting): Python ``bool``/``int``/``float`` -> bool, int32,
    float32; 64-bit numpy dtypes -> their 32-bit forms; unsigned 32-bit ->
    int32 storage (as ``uint32`` here); torch dtypes as they are."""
    if isinstance(dt, torch.dtype):
        return dt
    if dt is bool:
        return torch.bool
    if dt is int:
        return torch.int32
    if dt is float:
        return torch.float32
    nd = _np.dtype(dt)
    nd = {_np.dtype(_np.float64): _np.float32, _np.dtype(_np.int64): _np.int32,
          _np.dtype(_np.uint64): _np.int32, _np.dtype(_np.uint32): _np.int32}.get(nd, nd)
    return torch.from_numpy(_np.zeros(0, nd)).dtype


def _astype(self, dtype):
    """``jax.Array.astype``: a converted copy, dtype read as JAX does."""
    return self.to(jax_dtype(dtype))


# jax.Array.astype on torch tensors (reference model code calls it on state
# arrays, e.g. examples/models/advanced_economic_model.py:1493); torch has no
# method of that name, so nothing of torch's is replaced
if not hasattr(torch.Tensor, "astype"):
    torch.Tensor.astype = _astype


def asarray(a, dtype=None) -> torch.Tensor:
    """Array from a tensor, numpy array, Python scalar or (nested) sequence."""
    if dtype is not None:
        dtype = jax_dtype(dtype)
    if getattr(a, "_abm_sym", False) is True:
        return a.as_array() if dtype is None else a.as_array().to(dtype)
    tr = _trace()
    if tr is not None and _python_constant(a):
        # a constant inside a fused agent function: literals in the kernel
        out = tr.const(a).as_array() if isinstance(a, (bool, int, float)) else tr.stack(list(a))
        return out if dtype is None else out.to(dtype)
    if isinstance(a, torch.Tensor):
        return a if dtype is None else a.to(dtype)
    if isinstance(a, _np.ndarray):
        t = torch.from_numpy(_np.ascontiguousarray(a))
        if dtype is None and t.dtype == torch.float64:
            dtype = torch.float32
        if dtype is None and t.dtype == torch.int64:
            dtype = torch.int32
        return t.to(device=_dev(), dtype=dtype or t.dtype)
    if isinstance(a, (list, tuple)) and a and builtins.any(_traced(x) for x in a):
        tr = next(x for x in a if _traced(x)).tr
        out = tr.stack(a)      # a per-agent vector inside a fused agent function
        return out if dtype is None else out.to(dtype)
    if isinstance(a, (list, tuple)) and a and builtins.any(isinstance(x, torch.Tensor) for x in a):
        return torch.stack([asarray(x, dtype) for x in a])
    if isinstance(a, (bool, int, float)):
        # a fill kernel, not a host copy (no host synchronisation)
        return torch.full((), a, dtype=dtype or _default_dtype(a), device=_dev())
    if dtype is None:
        dtype = _default_dtype(a)
        if dtype is None:
            probe = _np.asarray(a)
            dtype = {"f": torch.float32, "i": torch.int32, "u": torch.int32,
                     "b": torch.bool}.get(probe.dtype.kind, torch.float32)
    return torch.as_tensor(a, dtype=dtype, device=_dev())


array = asarray


def _trace():
    from .fuse import active_trace

    return active_trace()


def _python_constant(a) -> bool:
    if isinstance(a, (bool, int, float)):
        return True
    return (isinstance(a, (list, tuple)) and 0 < len(a) <= 16 and
            builtins.all(isinstance(x, (bool, int, float)) for x in a))


def _traced(x) -> bool:
    return getattr(x, "_abm_sym", False) is True


def _pyscalar(x) -> bool:
    """A Python scalar, or a traced stand-in for one (env entries and the
    time step inside a fused agent function)."""
    return isinstance(x, (int, float)) or (_traced(x) and x.weak)


def _t(x):
    if isinstance(x, torch.Tensor):
        return x
    if _traced(x):
        return x.as_array()
    return asarray(x)


def zeros(shape, dtype=float32):
    return torch.zeros(shape, dtype=dtype, device=_dev())


def ones(shape, dtype=float32):
    return torch.ones(shape, dtype=dtype, device=_dev())


def full(shape, fill_value, dtype=None):
    return torch.full(shape if isinstance(shape, tuple) else (shape,) if isinstance(shape, int)
                      else tuple(shape), fill_value, dtype=dtype or _default_dtype(fill_value),
                      device=_dev())


def zeros_like(x, dtype=None):
    return torch.zeros_like(_t(x), dtype=dtype)


def ones_like(x, dtype=None):
    return torch.ones_like(_t(x), dtype=dtype)


def full_like(x, fill_value, dtype=None):
    return torch.full_like(_t(x), fill_value, dtype=dtype)


def arange(start, stop=None, step=1, dtype=None):
    if stop is None:
        start, stop = 0, start
    if dtype is None:
        dtype = torch.int32 if all(isinstance(v, int) for v in (start, stop, step)) else torch.float32
    return torch.arange(start, stop, step, dtype=dtype, device=_dev())


def linspace(start, stop, num=50, endpoint=True, dtype=float32):
    if endpoint:
        return torch.linspace(start, stop, num, dtype=dtype, device=_dev())
    return torch.linspace(start, stop, num + 1, dtype=dtype, device=_dev())[:-1]


def eye(n, dtype=float32):
    return torch.eye(n, dtype=dtype, device=_dev())


def where(cond, x=None, y=None):
    if x is None and y is None:
        return torch.nonzero(_t(cond), as_tuple=True)
    c = _t(cond)
    return torch.where(c, _t(x) if not _pyscalar(x) else x,
                       _t(y) if not _pyscalar(y) else y)


def clip(x, a_min=None, a_max=None):
    return torch.clamp(_t(x), min=a_min, max=a_max)


# ---------------------------------------------------------------- JAX semantics
_SUBTRACTIONS = frozenset({torch.sub, torch.subtract, torch.rsub, torch.Tensor.sub,
                           torch.Tensor.subtract, torch.Tensor.__sub__, torch.Tensor.__rsub__})
# torch's messages for what JaxSemantics does differently
JAX_SEMANTICS_ERRORS = ("with a bool tensor is not supported",
                        "calling .item() on a Tensor")


def _is_bool(x) -> bool:
    return isinstance(x, torch.Tensor) and x.dtype == torch.bool


def _int_tensor(x) -> bool:
    return isinstance(x, torch.Tensor) and not x.is_floating_point() and \
        not x.is_complex() and x.dtype != torch.bool


def _gather_rows(t: torch.Tensor, j: torch.Tensor) -> torch.Tensor:
    """t[j] along dim 0 as JAX's gather reads j: negative counts from the
    end, then clamped to [0, M-1]; a 0-d j selects one row without a host
    read (so it works under vmap and in captured graphs)."""
    m = t.shape[0]
    j = j.to(device=t.device, dtype=torch.long)
    j = torch.where(j < 0, j + m, j).clamp(0, m - 1)
    if j.dim() == 0:
        return t.index_select(0, j.reshape(1)).squeeze(0)
    return t[j]


class JaxSemantics(torch.overrides.TorchFunctionMode):
    """Two JAX behaviours torch lacks, for model code written against JAX:

    * a boolean operand of a subtraction is promoted (``1.0 - employed`` in
      the reference's examples/models/advanced_economic_model.py:184,
      ``x - mask``, ``mask - 1``) -- it is cast to torch's result type of the
      pair; bool - bool and -bool still raise, as they do in JAX;
    * indexing dim 0 with an integer tensor (``others[j]``, ``others[j, c]``,
      j per agent under vmap, where torch would read a 0-d j on the host)
      gathers with JAX's index rule: negative counts from the end, out of
      range clamps."""

    def __torch_function__(self, func, types, args=(), kwargs=None):
        kwargs = kwargs or {}
        if func in _SUBTRACTIONS and len(args) >= 2 and _is_bool(args[0]) != _is_bool(args[1]):
            dt = torch.result_type(args[0], args[1])
            if dt != torch.bool:
                args = tuple(a.to(dt) if _is_bool(a) else a for a in args[:2]) + tuple(args[2:])
        elif func is torch.Tensor.__getitem__ and len(args) == 2 and \
                isinstance(args[0], torch.Tensor) and args[0].dim() >= 1:
            t, idx = args
            if _int_tensor(idx):
                return _gather_rows(t, idx)
            if isinstance(idx, tuple) and idx and _int_tensor(idx[0]) and \
                    builtins.all(isinstance(i, (int, slice)) and not isinstance(i, bool)
                                 for i in idx[1:]):
                row = _gather_rows(t, idx[0])
                rest = idx[1:]
                if idx[0].dim() == 0:
                    return row[rest]
                return row[(slice(None),) * idx[0].dim() + rest]
        return func(*args, **kwargs)


def with_jax_semantics(fn):
    """``fn`` run inside :class:`JaxSemantics` (None stays None)."""
    if fn is None:
        return None

    def wrapped(*args, **kwargs):
        with JaxSemantics():
            return fn(*args, **kwargs)

    wrapped.__wrapped__ = fn
    return wrapped


def nan_to_num(x, copy=True, nan=0.0, posinf=None, neginf=None):
    """``jnp.nan_to_num``: NaN -> ``nan``, +inf -> ``posinf`` and -inf ->
    ``neginf`` (the infinities default to the dtype's largest finite
    magnitude); integer and boolean arrays come back unchanged (used by the
    reference's examples/models/advanced_economic_model.py:1860-1869)."""
    t = _t(x)
    if not t.is_floating_point():
        return t
    return torch.nan_to_num(t, nan=nan, posinf=posinf, neginf=neginf)


def _unary(fn):
    return lambda x: fn(_t(x))


abs = _unary(torch.abs)
absolute = abs
sign = _unary(torch.sign)
exp = _unary(torch.exp)
log = _unary(torch.log)
log1p = _unary(torch.log1p)
expm1 = _unary(torch.expm1)
sqrt = _unary(torch.sqrt)
square = _unary(torch.square)
sin = _unary(torch.sin)
cos = _unary(torch.cos)
tan = _unary(torch.tan)
tanh = _unary(torch.tanh)
arctan = _unary(torch.atan)
floor = _unary(torch.floor)
ceil = _unary(torch.ceil)
isnan = _unary(torch.isnan)
isfinite = _unary(torch.isfinite)
logical_not = _unary(torch.logical_not)
negative = _unary(torch.neg)


def round(x, decimals=0):
    return torch.round(_t(x), decimals=decimals)


def power(x, y):
    return torch.pow(_t(x), y)


def arctan2(y, x):
    return torch.atan2(_t(y), _t(x))


def mod(x, y):
    """Floored remainder with the sign of the divisor (jnp.mod)."""
    return torch.remainder(_t(x), y)


remainder = mod


def minimum(x, y):
    return torch.minimum(_t(x), _t(y))


def maximum(x, y):
    return torch.maximum(_t(x), _t(y))


def logical_and(x, y):
    return torch.logical_and(_t(x), _t(y))


def logical_or(x, y):
    return torch.logical_or(_t(x), _t(y))


def logical_xor(x, y):
    return torch.logical_xor(_t(x), _t(y))


def _reduce(fn, x, axis, keepdims, **kw):
    x = _t(x)
    if axis is None:
        if keepdims and isinstance(x, torch.Tensor) and x.dim() > 0:
            # JAX keeps every reduced axis as size 1
            return fn(x, dim=tuple(range(x.dim())), keepdim=True, **kw)
        return fn(x, **kw)
    return fn(x, dim=axis, keepdim=keepdims, **kw)


_SMALL_INTS = (torch.bool, torch.uint8, torch.int8, torch.int16, torch.int32)


def _int_acc(x):
    """JAX with x64 off (the reference's setting) sums and multiplies
    booleans and integers into int32 (wrapping), where torch would give
    int64."""
    return {"dtype": torch.int32} if x.dtype in _SMALL_INTS else {}


_DEVICE_REDUCIBLE = (torch.float32, torch.int32, torch.int16, torch.int8, torch.uint8, torch.bool)
_is_wrapped = torch._C._functorch.is_functorch_wrapped_tensor


def _reductions():
    from . import reductions

    return reductions


def _whole(x, axis) -> bool:
    """True when ``jnp.sum`` / ``jnp.mean`` of x over ``axis`` is a
    whole-array reduction of a plain GPU tensor that the HIP reduction
    kernels take: not under vmap or another functorch transform, not
    recording gradients."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda or x.numel() == 0:
        return False
    if axis is not None:
        axes = axis if isinstance(axis, (tuple, list)) else (axis,)
        rank = builtins.max(x.dim(), 1)
        if sorted(a % rank for a in axes) != list(range(x.dim())):
            return False
    if x.dtype not in _DEVICE_REDUCIBLE or x.requires_grad:
        return False
    return not _is_wrapped(x)


def sum(x, axis=None, keepdims=False):
    """jnp.sum.  A whole-array sum of a plain GPU tensor (the usual metrics
    and env-callback reduction) is one HIP launch (jaxabm_amd.reductions:
    float32 summed in float64, integers and booleans exactly into int32);
    other sums are torch reductions with JAX's x64-off dtypes."""
    x = _t(x)
    if not keepdims and _whole(x, axis):
        return _reductions().sum_all(x)
    return _reduce(torch.sum, x, axis, keepdims, **_int_acc(x))


def prod(x, axis=None, keepdims=False):
    x = _t(x)
    if axis is None:
        return torch.prod(x, **_int_acc(x))
    return torch.prod(x, dim=axis, keepdim=keepdims, **_int_acc(x))


def mean(x, axis=None, keepdims=False):
    """jnp.mean.  A whole-array mean of a plain GPU tensor is one HIP launch
    (float32 summed in float64; integers and booleans summed exactly, then
    float32(sum) / float32(n)); others are torch means in float32."""
    x = _t(x)
    if not keepdims and _whole(x, axis):
        return _reductions().mean_all(x)
    if not x.is_floating_point():
        x = x.to(torch.float32)
    return _reduce(torch.mean, x, axis, keepdims)


def var(x, axis=None, keepdims=False, ddof=0):
    x = _t(x).to(torch.float32) if not _t(x).is_floating_point() else _t(x)
    if axis is None:
        return torch.var(x, correction=ddof)
    return torch.var(x, dim=axis, keepdim=keepdims, correction=ddof)


def std(x, axis=None, keepdims=False, ddof=0):
    return torch.sqrt(var(x, axis, keepdims, ddof))


def min(x, axis=None, keepdims=False):
    x = _t(x)
    if axis is None:
        return torch.min(x)
    return torch.amin(x, dim=axis, keepdim=keepdims)


def max(x, axis=None, keepdims=False):
    x = _t(x)
    if axis is None:
        return torch.max(x)
    return torch.amax(x, dim=axis, keepdim=keepdims)


amin, amax = min, max


def argmin(x, axis=None):
    return torch.argmin(_t(x), dim=axis).to(torch.int32)


def argmax(x, axis=None):
    return torch.argmax(_t(x), dim=axis).to(torch.int32)


def any(x, axis=None, keepdims=False):
    x = _t(x).bool()
    if axis is None:
        return torch.any(x)
    return torch.any(x, dim=axis, keepdim=keepdims)


def all(x, axis=None, keepdims=False):
    x = _t(x).bool()
    if axis is None:
        return torch.all(x)
    return torch.all(x, dim=axis, keepdim=keepdims)


def count_nonzero(x, axis=None):
    x = _t(x)
    if _whole(x, axis):
        return _reductions().sum_all(x if x.dtype == torch.bool else x != 0)
    return torch.count_nonzero(x, dim=axis).to(torch.int32)


def cumsum(x, axis=None):
    x = _t(x)
    if axis is None:
        x, axis = x.reshape(-1), 0
    return torch.cumsum(x, dim=axis, **_int_acc(x))


def diff(x, axis=-1):
    return torch.diff(_t(x), dim=axis)


def sort(x, axis=-1):
    return torch.sort(_t(x), dim=axis, stable=True).values


def argsort(x, axis=-1):
    return torch.argsort(_t(x), dim=axis, stable=True).to(torch.int32)


def unique(x):
    return torch.unique(_t(x), sorted=True)


def bincount(x, weights=None, minlength=0, length=None):
    x = _t(x).long()
    n = length if length is not None else minlength
    out = torch.bincount(x, weights=None if weights is None else _t(weights), minlength=n)
    if length is not None:
        out = out[:length]
    return out.to(torch.int32) if weights is None else out


def stack(arrays, axis=0):
    return torch.stack([_t(a) for a in arrays], dim=axis)


def concatenate(arrays, axis=0):
    return torch.cat([_t(a) for a in arrays], dim=axis)


def column_stack(arrays):
    return torch.stack([_t(a) for a in arrays], dim=1)


def reshape(x, shape):
    return _t(x).reshape(shape)


def transpose(x, axes=None):
    x = _t(x)
    return x.permute(*axes) if axes is not None else x.permute(*reversed(range(x.dim())))


def expand_dims(x, axis):
    return _t(x).unsqueeze(axis)


def squeeze(x, axis=None):
    return _t(x).squeeze() if axis is None else _t(x).squeeze(axis)


def _take_fill(dtype):
    """jnp.take's default fill value for out-of-range indices: NaN for
    floats, the most negative value for signed integers, the largest for
    unsigned ones, True for booleans."""
    if dtype.is_floating_point or dtype.is_complex:
        return float("nan")
    if dtype == torch.bool:
        return True
    return torch.iinfo(dtype).min if dtype.is_signed else torch.iinfo(dtype).max


def take(x, indices, axis=None):
    """jnp.take with its default mode "fill": negative indices count from the
    end, and an index still outside the axis yields the fill value (NaN,
    the most negative integer, the largest unsigned, True)."""
    if getattr(x, "_abm_sym", False) and isinstance(indices, int) and \
            not isinstance(indices, bool) and axis in (None, 0, -1) and hasattr(x, "items"):
        # a per-agent vector inside a traced agent function, a constant index
        k = len(x)
        i = indices + k if indices < 0 else indices
        if not 0 <= i < k:
            from .fuse import Unsupported
            raise Unsupported("jnp.take out of range (the fill value) in a traced function")
        return x[i]
    x = _t(x)
    idx = _t(indices).to(device=x.device, dtype=torch.long)
    src = x.reshape(-1) if axis is None else x
    ax = 0 if axis is None else axis % x.dim()
    size = src.shape[ax]
    idx = torch.where(idx < 0, idx + size, idx)
    valid = (idx >= 0) & (idx < size)
    safe = idx.clamp(0, builtins.max(size - 1, 0))
    if axis is None:
        out = src[safe] if size else torch.zeros(idx.shape, dtype=x.dtype, device=x.device)
        return torch.where(valid, out, torch.full_like(out, _take_fill(x.dtype)))
    out = torch.index_select(src, ax, safe.reshape(-1)).reshape(
        *x.shape[:ax], *idx.shape, *x.shape[ax + 1:])
    mask = valid.reshape((1,) * ax + tuple(idx.shape) + (1,) * (x.dim() - ax - 1))
    return torch.where(mask, out, torch.full_like(out, _take_fill(x.dtype)))


def flip(x, axis=None):
    """jnp.flip: reverse the order along ``axis`` (every axis when None)."""
    x = _t(x)
    dims = tuple(range(x.dim())) if axis is None else \
        (tuple(axis) if isinstance(axis, (tuple, list)) else (axis,))
    return torch.flip(x, dims)


def roll(x, shift, axis=None):
    """jnp.roll: shift along ``axis``; with no axis the flattened array is
    rolled and the shape kept."""
    x = _t(x)
    if axis is None:
        return torch.roll(x, shift)
    return torch.roll(x, shift, axis)


def median(x, axis=None, keepdims=False):
    """jnp.median: the middle value, the mean of the two middle values for
    an even count (numpy's rule); integers promote to float32; a NaN gives
    NaN."""
    x = _t(x)
    if not x.is_floating_point():
        x = x.to(torch.float32)
    if axis is None:
        return torch.quantile(x.reshape(-1), 0.5)
    return torch.quantile(x, 0.5, dim=axis, keepdim=keepdims)


def hypot(a, b):
    """jnp.hypot: sqrt(a^2 + b^2) without overflow; integers promote to
    float32."""
    a, b = _t(a), _t(b)
    if not a.is_floating_point():
        a = a.to(torch.float32)
    if not b.is_floating_point():
        b = b.to(torch.float32)
    return torch.hypot(a, b)


def dot(a, b):
    return torch.matmul(_t(a), _t(b))


matmul = dot


def outer(a, b):
    return torch.outer(_t(a), _t(b))


def allclose(a, b, rtol=1e-5, atol=1e-8):
    return builtins.bool(torch.allclose(_t(a).to(torch.float64), _t(b).to(torch.float64),
                                        rtol=rtol, atol=atol))


def array_equal(a, b):
    a, b = _t(a), _t(b)
    return builtins.bool(a.shape == b.shape and torch.equal(a, b.to(a.dtype)))


def _int_parts(y: torch.Tensor, index):
    """An integer index of y's leading dims (an integer tensor, a Python int,
    or a tuple of those), as JAX's scatter and gather read it: each part
    normalised (a negative value counts from the end) and broadcast, with the
    mask of the entries inside y.  None for other indices (slices, masks,
    Ellipsis), which torch's indexing handles."""
    parts = index if isinstance(index, tuple) else (index,)
    if not parts or len(parts) > y.dim():
        return None
    if not builtins.all(_int_index(p) or (isinstance(p, int) and not isinstance(p, bool))
                        for p in parts):
        return None
    idx, valid = [], None
    for d, p in enumerate(parts):
        t = torch.as_tensor(p, device=y.device).to(torch.long)
        size = y.shape[d]
        t = torch.where(t < 0, t + size, t)
        ok = (t >= 0) & (t < size)
        idx.append(t)
        valid = ok if valid is None else valid & ok
    idx = list(torch.broadcast_tensors(*idx))
    return idx, valid.expand(idx[0].shape)


def _scatter_dropping(y: torch.Tensor, index, update):
    """Applies ``update(target, idx)`` (an in-place scatter) with JAX's rule
    for updates outside the array: they are dropped.  Such entries are sent
    to one spare row past the end of each indexed dim, which is cut off
    afterwards (no host read of the indices, so vmap and graph capture keep
    working).  Returns the updated copy, or None for an index this does not
    cover."""
    r = _int_parts(y, index)
    if r is None:
        return None
    idx, valid = r
    k = len(idx)
    inner = tuple(slice(0, sz) for sz in y.shape[:k])
    ext = y.new_zeros(tuple(sz + 1 for sz in y.shape[:k]) + tuple(y.shape[k:]))
    ext[inner] = y
    idx = tuple(torch.where(valid, t, torch.full_like(t, y.shape[d])) for d, t in enumerate(idx))
    update(ext, idx)
    return ext[inner].clone()


def at_set(x, index, value):
    """``x.at[index].set(value)`` (out of place); an integer index outside
    x is dropped, as JAX's scatter does."""
    y = _t(x)
    v = value.to(y.dtype) if isinstance(value, torch.Tensor) else \
        torch.as_tensor(value, dtype=y.dtype, device=y.device)

    def put(t, idx):
        t[idx] = v

    out = _scatter_dropping(y, index, put)
    if out is not None:
        return out
    y = y.clone()
    y[index] = value
    return y


def at_add(x, index, value):
    """``x.at[index].add(value)`` (out of place): the value is converted to
    x's dtype; repeated integer indices accumulate; indices outside x are
    dropped."""
    y = _t(x)
    v = value.to(y.dtype) if isinstance(value, torch.Tensor) else \
        torch.as_tensor(value, dtype=y.dtype, device=y.device)
    if isinstance(index, list):
        index = torch.as_tensor(index, device=y.device)

    def acc(t, idx):
        t.index_put_(idx, v.expand_as(t[idx]), accumulate=True)

    out = _scatter_dropping(y, index, acc)
    if out is not None:
        return out
    y = y.clone()
    y[index] = y[index] + v
    return y


def _int_index(index) -> bool:
    return isinstance(index, torch.Tensor) and not index.is_floating_point() and \
        index.dtype != torch.bool


class _AtRef:
    """``x.at[index]``: out-of-place updates with JAX's meaning
    (jax.numpy.ndarray.at).  Repeated integer indices accumulate in ``add``,
    ``multiply``, ``min`` and ``max``; ``set`` keeps one of the values."""

    __slots__ = ("x", "index")

    def __init__(self, x, index):
        self.x, self.index = x, index

    def get(self):
        """Gather with JAX's rule: negative indices count from the end and
        an index still outside is clamped."""
        r = _int_parts(self.x, self.index)
        if r is None:
            return self.x[self.index]
        idx, _ = r
        return self.x[tuple(t.clamp(0, builtins.max(self.x.shape[d] - 1, 0))
                            for d, t in enumerate(idx))]

    def set(self, value):
        return at_set(self.x, self.index, value)

    def add(self, value):
        return at_add(self.x, self.index, value)

    def _reduce(self, value, how: str, elementwise):
        y = self.x
        v = value.to(y.dtype) if isinstance(value, torch.Tensor) else \
            torch.as_tensor(value, dtype=y.dtype, device=y.device)
        if _int_index(self.index) and self.index.dim() == 1:
            def red(t, idx):
                src = v.expand((idx[0].shape[0],) + tuple(t.shape[1:]))
                with warnings.catch_warnings():   # torch marks index_reduce as beta
                    warnings.simplefilter("ignore", UserWarning)
                    t.index_reduce_(0, idx[0], src.contiguous(), how, include_self=True)
        else:
            def red(t, idx):
                t[idx] = elementwise(t[idx], v)
        out = _scatter_dropping(y, self.index, red)
        if out is not None:
            return out
        y = y.clone()
        y[self.index] = elementwise(y[self.index], v)
        return y

    def multiply(self, value):
        return self._reduce(value, "prod", torch.mul)

    mul = multiply

    def divide(self, value):
        """x / (product of the values at each index), as JAX computes it."""
        ones = torch.ones_like(self.x)
        den = _AtRef(ones, self.index).multiply(value)
        return self.x / den

    def power(self, value):
        def pw(t, idx):
            t[idx] = t[idx] ** value

        out = _scatter_dropping(self.x, self.index, pw)
        if out is not None:
            return out
        y = self.x.clone()
        y[self.index] = y[self.index] ** value
        return y

    def min(self, value):
        return self._reduce(value, "amin", torch.minimum)

    def max(self, value):
        return self._reduce(value, "amax", torch.maximum)

    def apply(self, fn):
        def ap(t, idx):
            t[idx] = fn(t[idx])

        out = _scatter_dropping(self.x, self.index, ap)
        if out is not None:
            return out
        y = self.x.clone()
        y[self.index] = fn(y[self.index])
        return y


class _AtIndexer:
    __slots__ = ("x",)

    def __init__(self, x):
        self.x = x

    def __getitem__(self, index):
        return _AtRef(self.x, index)


# jax.Array.at on torch tensors (torch has no attribute of that name)
if not hasattr(torch.Tensor, "at"):
    torch.Tensor.at = property(lambda self: _AtIndexer(self))


class linalg:  # noqa: N801 -- mirrors jnp.linalg
    @staticmethod
    def norm(x, ord=None, axis=None, keepdims=False):
        x = _t(x)
        if not x.is_floating_point():
            x = x.to(torch.float32)
        return torch.linalg.norm(x, ord=ord, dim=axis, keepdim=keepdims)
