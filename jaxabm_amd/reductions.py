"""Record-time reductions over one agent variable on the GPU: state counts,
mean, Gini and the soft histogram.

Replaces the metric reductions a reference model computes each step with
jnp (jaxabm/model.py:194-200 metrics_fn, jaxabm/agentpy.py:926-942
compute_metrics, jaxabm/agent.py:198-211 aggregate) and the Gini definition
of examples/models/advanced_economic_model.py:1798-1811.  Results are device
tensors: nothing waits on the host, and the kernels clear their outputs and
scratch with fill kernels (no memset), so a step that uses them can be
replayed as a graph (``Model.run(graph=True)``).

* ``state_counts(s, k)``: int32 counts of each value < k of a uint8 array,
  exact.
* ``mean(x)``: int32 input as the reference's jnp.mean of int32 (exact sum,
  then float32(sum) / float32(n)); float32 input summed in float64 in a fixed
  order and rounded once (deterministic; the reference's float32 summation
  order is XLA's, so float means agree to rounding).
* ``sum_all(x)`` / ``mean_all(x)``: the whole-array jnp.sum / jnp.mean that
  ``jaxabm_amd.numpy`` runs on plain GPU tensors (env and metrics callbacks):
  one launch up to 65,536 values.  Inside ``deferring()`` (a model step's
  env and metrics callbacks) those launches wait and go out together, one
  ``abm_reduce_batch`` launch, when the scope ends or when anything reads
  one of their results, whichever comes first; each value is computed
  exactly as its own launch computes it.
* ``gini(x)``: int32 input, the reference's formula bit for bit (int32 sums
  with wrap-around, float32 quotient).
* ``soft_histogram(x, bins, slope)``: float64 bin means of sigmoid
  differences with edges k / bins (the Deffuant calibration statistic).
"""
from __future__ import annotations

import contextlib
import threading

import torch

from ._lib import (REDUCE_BATCH_MAX, REDUCE_MEAN_F32, REDUCE_MEAN_I32, REDUCE_SUM_F32,
                   REDUCE_SUM_I32, REDUCE_SUM_U8, ReduceItem, check, lib, ptr, require_gpu,
                   stream_handle)



def _workspace(n: int, device) -> torch.Tensor:
    nbytes = int(lib().abm_reduce_workspace(int(n)))
    return torch.empty(nbytes, dtype=torch.uint8, device=device)


def _check_input(x: torch.Tensor, dtypes) -> torch.Tensor:
    require_gpu()
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("reductions take a GPU tensor")
    if x.dtype not in dtypes:
        raise TypeError(f"dtype {x.dtype} not supported here (expected one of {dtypes})")
    x = x.reshape(-1)
    return x if x.is_contiguous() else x.contiguous()


def state_counts(states: torch.Tensor, k: int) -> torch.Tensor:
    """int32[k]: how many entries of ``states`` (uint8) equal 0, 1, ..., k-1."""
    s = _check_input(states, (torch.uint8,))
    out = torch.empty(int(k), dtype=torch.int32, device=s.device)
    check(lib().abm_reduce_counts_u8(ptr(s), s.numel(), int(k), ptr(out), stream_handle()),
          "reduce_counts_u8")
    return out


def mean(x: torch.Tensor) -> torch.Tensor:
    """0-d float32 mean of an int32 or float32 array."""
    _check_input(x, (torch.int32, torch.float32))
    return mean_all(x)


_SUM_SMALL = 65536   # abm_sum_* / abm_mean_* need no workspace up to this size

# whole-array reductions are called from model callbacks every step: the
# entry points and dtypes are looked up once, and the stream is read as a
# raw handle
_FNS: dict = {}


def _fn(name: str):
    f = _FNS.get(name)
    if f is None:
        f = _FNS[name] = getattr(lib(), name)
    return f


# ---- batched launches of the step's whole-array reductions
_tls = threading.local()
_OPS = {"abm_sum_f32": REDUCE_SUM_F32, "abm_mean_f32": REDUCE_MEAN_F32,
        "abm_sum_i32": REDUCE_SUM_I32, "abm_mean_i32": REDUCE_MEAN_I32,
        "abm_sum_u8": REDUCE_SUM_U8}
_METADATA = {torch.Tensor.dtype.__get__, torch.Tensor.shape.__get__,
             torch.Tensor.device.__get__, torch.Tensor.is_cuda.__get__,
             torch.Tensor.requires_grad.__get__, torch.Tensor.ndim.__get__, torch.Tensor.dim,
             torch.Tensor.numel, torch.Tensor.size, torch.Tensor.is_floating_point}


def _plain(obj):
    if isinstance(obj, _Deferred):
        return obj._abm_plain
    if isinstance(obj, (list, tuple)):
        return type(obj)(_plain(o) for o in obj)
    if isinstance(obj, dict):
        return {k: _plain(v) for k, v in obj.items()}
    return obj


class _Deferred(torch.Tensor):
    """The 0-d result of a whole-array sum or mean whose launch waits for
    the rest of the step's reductions (``deferring()``).  Any torch operation
    on it launches the waiting batch first and then runs on the plain
    tensor; reading its dtype, shape or device does not."""

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        if func not in _METADATA:
            flush()
        return func(*_plain(args), **_plain(kwargs or {}))


def _pending() -> list:
    p = getattr(_tls, "pending", None)
    if p is None:
        p = _tls.pending = []
    return p


def flush() -> None:
    """Launch the waiting reductions (abm_reduce_batch, up to 16 a launch)."""
    pend = _pending()
    if not pend:
        return
    _tls.pending = []
    for _op, v, _out, _s, version in pend:
        if v._version != version:
            # the eager launch would have read the values before the change
            raise RuntimeError("an array was modified in place after a whole-array sum or mean "
                               "of it inside a model callback; compute the reduction after the "
                               "update, or update a copy")
    stream = pend[0][3]
    for i in range(0, len(pend), REDUCE_BATCH_MAX):
        part = pend[i:i + REDUCE_BATCH_MAX]
        items = (ReduceItem * len(part))(*[ReduceItem(v.data_ptr(), v.numel(), out.data_ptr(),
                                                      op, 0) for op, v, out, _s, _ver in part])
        check(_fn("abm_reduce_batch")(items, len(part), stream), "reduce_batch")


@contextlib.contextmanager
def deferring():
    """Batch the whole-array sums and means issued inside into one launch
    (at the end of the outermost scope, or earlier when a result is read)."""
    depth = getattr(_tls, "depth", 0)
    _tls.depth = depth + 1
    try:
        yield
    finally:
        _tls.depth = depth
        if depth == 0:
            flush()


def materialize(d):
    """``d`` (a dict of callback results) with deferred results replaced by
    plain tensors, the waiting reductions launched first."""
    flush()
    if isinstance(d, dict) and any(isinstance(v, _Deferred) for v in d.values()):
        return {k: (v._abm_plain if isinstance(v, _Deferred) else v) for k, v in d.items()}
    return d


def _whole_call(name: str, v: torch.Tensor, out_dtype) -> torch.Tensor:
    n = v.numel()
    dev = v.device
    out = torch.empty((), dtype=out_dtype, device=dev)
    stream = torch._C._cuda_getCurrentRawStream(dev.index if dev.index is not None
                                                else torch.cuda.current_device())
    if n <= _SUM_SMALL and getattr(_tls, "depth", 0):
        pend = _pending()
        if pend and pend[0][3] != stream:
            flush()
        # the input stays referenced until its launch is enqueued, and must
        # not change before then (checked by its version counter)
        _pending().append((_OPS[name], v, out, stream, v._version))
        d = torch.Tensor._make_subclass(_Deferred, out)
        d._abm_plain = out
        return d
    ws = None if n <= _SUM_SMALL else _workspace(n, dev)
    rc = _fn(name)(v.data_ptr(), n, out.data_ptr(), None if ws is None else ws.data_ptr(),
                   0 if ws is None else ws.numel(), stream)
    if rc:
        check(rc, name)
    return out


def sum_all(x: torch.Tensor) -> torch.Tensor:
    """0-d sum of every element, as jnp.sum with x64 off: float32 (summed in
    float64 in a fixed order, rounded once), or int32 with wrap-around for
    int32, int16, int8, uint8 and bool input."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("reductions take a GPU tensor")
    dt = x.dtype
    if dt in (torch.int8, torch.int16):
        x, dt = x.to(torch.int32), torch.int32
    if dt not in (torch.float32, torch.int32, torch.uint8, torch.bool):
        raise TypeError(f"dtype {dt} not supported here")
    v = x.reshape(-1)
    if not v.is_contiguous():
        v = v.contiguous()
    if v.numel() == 0:
        return torch.zeros((), dtype=torch.float32 if dt == torch.float32 else torch.int32,
                           device=v.device)
    if dt == torch.float32:
        return _whole_call("abm_sum_f32", v, torch.float32)
    if dt == torch.int32:
        return _whole_call("abm_sum_i32", v, torch.int32)
    return _whole_call("abm_sum_u8", v.view(torch.uint8) if dt == torch.bool else v, torch.int32)


def mean_all(x: torch.Tensor) -> torch.Tensor:
    """0-d float32 mean of every element: float32 input summed in float64
    (fixed order), integer and boolean input summed exactly, then
    float32(sum) / float32(n)."""
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise TypeError("reductions take a GPU tensor")
    if x.dtype in (torch.int8, torch.int16, torch.uint8, torch.bool):
        x = x.to(torch.int32)
    if x.dtype not in (torch.int32, torch.float32):
        raise TypeError(f"dtype {x.dtype} not supported here")
    v = x.reshape(-1)
    if not v.is_contiguous():
        v = v.contiguous()
    if v.numel() == 0:
        raise ValueError("mean of an empty array")
    return _whole_call("abm_mean_i32" if v.dtype == torch.int32 else "abm_mean_f32", v,
                       torch.float32)


def gini(x: torch.Tensor) -> torch.Tensor:
    """0-d float32 Gini coefficient of an int32 array (the reference's
    definition: 2 sum_i i x_(i) / (n sum x) - (n + 1) / n, 0 when the sum is
    not positive)."""
    v = _check_input(x, (torch.int32,))
    if v.numel() == 0:
        raise ValueError("gini of an empty array")
    out = torch.empty((), dtype=torch.float32, device=v.device)
    ws = _workspace(v.numel(), v.device)
    check(lib().abm_gini_i32(ptr(v), v.numel(), ptr(out), ptr(ws), ws.numel(), stream_handle()),
          "gini_i32")
    return out


def soft_histogram(x: torch.Tensor, bins: int = 32, slope: float = 128.0) -> torch.Tensor:
    """float64[bins]: mean over the values of sigmoid(slope (x - k/bins)) -
    sigmoid(slope (x - (k+1)/bins))."""
    v = _check_input(x, (torch.float32,))
    if v.numel() == 0:
        raise ValueError("soft histogram of an empty array")
    out = torch.empty(int(bins), dtype=torch.float64, device=v.device)
    ws = _workspace(v.numel(), v.device)
    check(lib().abm_soft_hist_f32(ptr(v), v.numel(), int(bins), float(slope), ptr(out), ptr(ws),
                                  ws.numel(), stream_handle()), "soft_hist_f32")
    return out
