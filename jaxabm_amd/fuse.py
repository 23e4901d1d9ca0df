"""Per-agent functions traced into one generated HIP kernel.

The reference maps ``AgentType.init_state`` / ``update`` over the agents with
``jax.vmap`` inside ``AgentCollection.init`` / ``update``
(jaxabm/agent.py:92-177), and XLA fuses the mapped program.  Run through
``torch.func.vmap`` instead, every primitive of the function is a separate
kernel over all agents (a 6-op update with two draws is about 20 launches).
Here the function is run once on symbolic per-agent values, the operations
are recorded, and one kernel is generated in which thread i computes agent
i's whole update.  That kernel is compiled at run time with hiprtc
(``abm_rtc_compile``), cached by its source text, and launched on the
current stream, so it is also captured by ``Model.run(graph=True)``.

Semantics are those of the vmapped torch program, bit for bit except the
sign bit of a NaN result: torch subtracts as ``a + (-1) * b`` with a run-time
multiply, the kernel as ``a - b``, whose negate modifier flips a NaN
operand's sign.

* Result dtypes come from ``torch.result_type`` on stand-ins with the same
  roles: a per-agent value is a 1-D tensor (it is batched under vmap), a 0-d
  env tensor is 0-d, and a Python scalar (env entries, ``time_step``,
  literals) is a weak Python scalar.  Arithmetic between Python scalars is
  done as Python does it (float64, int64 with floored ``//`` and ``%``).
* Every operation rounds where the torch op would: the kernel is compiled
  without contraction into fused multiply-adds, literals keep their exact
  float64 bits and are converted to float32 where torch converts them, and
  ``remainder``, ``floor_divide``, ``clamp``, ``minimum``, ``maximum`` and
  ``sign`` follow torch's element functions.  A tensor divided by a Python
  scalar is multiplied by the scalar's reciprocal and a Python scalar
  divided by a tensor is ``reciprocal(tensor) * scalar``, as torch computes
  them.
* Draws are the ``abm_device.hpp`` functions the vmapped draw kernels call:
  agent i's key is ``split(key, n)[i]``; ``split``, ``bits``, ``uniform``,
  ``normal``, ``bernoulli``, ``randint``, ``gamma``, ``loggamma`` and
  ``beta`` on it give the same values as the vmapped draws (``lognormal``
  and ``exponential`` are built from ``normal`` and ``uniform``).
* Per-agent vectors of up to 16 values are unrolled.  float32 sums, means,
  products, 2-norms, variances and standard deviations over 1 to 16 of them
  follow the order torch's reduction uses for a row of that length, and dot
  products of two such vectors of up to 15 values the order of torch's
  batched product; cumulative sums follow torch's scan, and stable sorts
  and argsorts torch's radix key order; integer sums, products and
  cumulative sums are exact.

* ``model_state['agents_<name>'][var][j]`` (another agent's value, j a
  per-agent integer) is a gather; out-of-range indices follow JAX's clamp.
  Whole-array use of those arrays (``jnp.mean(...)``, a constant index)
  gives the same value to every agent and runs as torch ops before the
  kernel, as it does once under vmap.

Anything else is not traced.  That includes longer vectors and other
reductions, Python control flow on a traced value
and host reads; ``Unsupported`` is raised, and the collection runs the
vmapped path, which is remembered per function and input signature.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import math
import os
import struct
from typing import Any, Dict, List, Optional, Tuple

import numpy as np
import torch

_ENABLED = True
_ARCH = os.environ.get("JAXABM_AMD_ARCH", "gfx950")
_CSRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc")
# the generated kernels are compiled from the text of these files, in order
# (abm_device.hpp includes abm_math.h only when not compiled by hiprtc)
_PRELUDE_PATHS = (os.path.join(_CSRC, "abm_math.h"), os.path.join(_CSRC, "abm_device.hpp"))


def enable(flag: bool = True) -> None:
    """Turn the fused path on or off (on by default)."""
    global _ENABLED
    _ENABLED = bool(flag)


def enabled() -> bool:
    return _ENABLED


class Unsupported(Exception):
    """The function uses something the generated kernel does not cover."""


# kinds: tensor dtypes, and the three Python scalar types (weak)
_CT = {"f32": "float", "f64": "double", "i32": "int32_t", "i64": "int64_t", "u8": "uint8_t",
       "bool": "bool", "pyf": "double", "pyi": "int64_t", "pyb": "bool"}
_TORCH = {"f32": torch.float32, "f64": torch.float64, "i32": torch.int32, "i64": torch.int64,
          "u8": torch.uint8, "bool": torch.bool}
_KIND = {v: k for k, v in _TORCH.items()}
_WEAK = ("pyf", "pyi", "pyb")
_PROMOTED: Dict[tuple, str] = {}   # (kind a, varying a, kind b, varying b) -> result kind
_FLOATS = ("f32", "f64", "pyf")
_DEFAULT = {"pyf": "f32", "pyi": "i32", "pyb": "bool"}   # JAX's defaults for Python scalars


def _lit(v, kind: str) -> str:
    """C expression of the Python value v converted to kind, as torch
    converts a Python scalar operand (through float64 for floats)."""
    ct = _CT[kind]
    if isinstance(v, bool):
        return f"(({ct}){'true' if v else 'false'})"
    if isinstance(v, int):
        if not -(1 << 63) <= v < (1 << 63):
            raise Unsupported("integer literal out of int64 range")
        return f"(({ct})({v}LL))" if v != -(1 << 63) else f"(({ct})(-9223372036854775807LL - 1))"
    if isinstance(v, float):
        if math.isnan(v):
            return f"(({ct})__longlong_as_double(0x7ff8000000000000LL))"
        if math.isinf(v):
            return f"(({ct})__longlong_as_double({'0x7ff0000000000000LL' if v > 0 else '(long long)0xfff0000000000000ULL'}))"
        return f"(({ct}){float(v).hex()})"
    raise Unsupported(f"literal of type {type(v).__name__}")


class Sym:
    """A traced value: one scalar per agent (varying) or one for all agents
    (uniform: an env value, a literal, a draw-free expression of those)."""

    _abm_sym = True
    __slots__ = ("tr", "name", "kind", "varying")

    def __init__(self, tr: "_Trace", name: str, kind: str, varying: bool):
        self.tr, self.name, self.kind, self.varying = tr, name, kind, varying

    # -- what torch sees
    @property
    def weak(self) -> bool:
        return self.kind in _WEAK

    @property
    def dtype(self):
        if self.weak:
            raise Unsupported("dtype of a Python scalar")
        return _TORCH[self.kind]

    @property
    def shape(self):
        return torch.Size(())

    def dim(self):
        return 0

    def size(self, *a):
        return torch.Size(())

    def is_floating_point(self):
        return self.kind in _FLOATS

    def as_array(self) -> "Sym":
        """jnp.asarray of a Python scalar: a 0-d tensor of JAX's default dtype."""
        if not self.weak:
            return self
        return self.tr.cast(self, _DEFAULT[self.kind])

    def to(self, *args, **kwargs):
        dt = kwargs.get("dtype")
        for a in args:
            if isinstance(a, torch.dtype):
                dt = a
        if dt is None:
            return self
        if dt not in _KIND:
            raise Unsupported(f"cast to {dt}")
        return self.tr.cast(self, _KIND[dt])

    def astype(self, dt):
        from .numpy import jax_dtype

        return self.to(jax_dtype(dt))

    def float(self):
        return self.to(torch.float32)

    def double(self):
        return self.to(torch.float64)

    def int(self):
        return self.to(torch.int32)

    def long(self):
        return self.to(torch.int64)

    def bool(self):
        return self.to(torch.bool)

    def clone(self):
        return self

    def contiguous(self):
        return self

    def detach(self):
        return self

    def reshape(self, *shape):
        if shape in ((), ((),)) or shape == ((1,),) or shape == (1,):
            return self
        raise Unsupported("reshape of a per-agent scalar")

    def __bool__(self):
        raise Unsupported("Python control flow on a traced value")

    def __int__(self):
        raise Unsupported("int() of a traced value")

    def __float__(self):
        raise Unsupported("float() of a traced value")

    def __index__(self):
        raise Unsupported("indexing with a traced value")

    def item(self):
        raise Unsupported("item() of a traced value")

    def __len__(self):
        raise Unsupported("len() of a per-agent scalar")

    def __iter__(self):
        raise Unsupported("iteration over a per-agent scalar")

    def __getitem__(self, i):
        raise Unsupported("indexing a per-agent scalar")

    # -- operators
    def __add__(self, o): return self.tr.binary("add", self, o)
    def __radd__(self, o): return self.tr.binary("add", o, self)
    def __sub__(self, o): return self.tr.binary("sub", self, o)
    def __rsub__(self, o): return self.tr.binary("sub", o, self)
    def __mul__(self, o): return self.tr.binary("mul", self, o)
    def __rmul__(self, o): return self.tr.binary("mul", o, self)
    def __truediv__(self, o): return self.tr.binary("div", self, o)
    def __rtruediv__(self, o): return self.tr.binary("div", o, self)
    def __floordiv__(self, o): return self.tr.binary("floordiv", self, o)
    def __rfloordiv__(self, o): return self.tr.binary("floordiv", o, self)
    def __mod__(self, o): return self.tr.binary("mod", self, o)
    def __rmod__(self, o): return self.tr.binary("mod", o, self)
    def __pow__(self, o): return self.tr.power(self, o)
    def __lt__(self, o): return self.tr.binary("lt", self, o)
    def __le__(self, o): return self.tr.binary("le", self, o)
    def __gt__(self, o): return self.tr.binary("gt", self, o)
    def __ge__(self, o): return self.tr.binary("ge", self, o)
    def __eq__(self, o): return self.tr.binary("eq", self, o)  # noqa: PLE0307
    def __ne__(self, o): return self.tr.binary("ne", self, o)
    def __and__(self, o): return self.tr.binary("and", self, o)
    def __rand__(self, o): return self.tr.binary("and", o, self)
    def __or__(self, o): return self.tr.binary("or", self, o)
    def __ror__(self, o): return self.tr.binary("or", o, self)
    def __xor__(self, o): return self.tr.binary("xor", self, o)
    def __rxor__(self, o): return self.tr.binary("xor", o, self)
    def __neg__(self): return self.tr.unary("neg", self)
    def __pos__(self): return self
    def __abs__(self): return self.tr.unary("abs", self)
    def __invert__(self): return self.tr.unary("invert", self)

    __hash__ = object.__hash__

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        kwargs = kwargs or {}
        tr = None
        for a in list(args) + list(kwargs.values()):
            if isinstance(a, (Sym, SymVec)):
                tr = a.tr
                break
            if isinstance(a, (list, tuple)):
                for b in a:
                    if isinstance(b, (Sym, SymVec)):
                        tr = b.tr
        if tr is None:
            raise Unsupported(f"{getattr(func, '__name__', func)}")
        return tr.torch_call(func, args, kwargs)


class SymKey:
    """A traced PRNG key (two uint32 words per agent)."""

    _abm_symkey = True
    __slots__ = ("tr", "name")

    def __init__(self, tr: "_Trace", name: str):
        self.tr, self.name = tr, name

    @property
    def shape(self):
        return torch.Size((2,))

    def _bits(self, per: int = 1, j: int = 0) -> str:
        if per == 1:
            return self.tr.tmp("uint32_t", f"abm::bits_scalar({self.name}, {self.tr.mode})")
        return self.tr.tmp("uint32_t", f"abm::bits_at({self.name}, {per}LL, {j}LL, {self.tr.mode})")

    def _shaped(self, shape, one):
        """one(per, j) for every element of a draw of ``shape`` (() or (k,))."""
        shape = tuple(shape)
        if shape == ():
            return one(1, 0)
        if len(shape) != 1 or not 0 < shape[0] <= _MAX_VEC:
            raise Unsupported("a draw of more than one dimension or 16 elements")
        return SymVec(self.tr, [one(shape[0], j) for j in range(shape[0])])

    def split(self, num: int) -> "SymKeyArray":
        num = int(num)
        if num < 1:
            raise Unsupported("split into fewer than one key")
        tr = self.tr
        if num == 2:
            a, b = tr.fresh("k"), tr.fresh("k")
            tr.lines.append(f"abm::Key {a}, {b}; abm::split2({self.name}, {a}, {b}, {tr.mode});")
            return SymKeyArray([SymKey(tr, a), SymKey(tr, b)])
        out = []
        for j in range(num):
            k = tr.fresh("k")
            tr.lines.append(f"const abm::Key {k} = abm::split_at({self.name}, {num}LL, {j}LL, "
                            f"{tr.mode});")
            out.append(SymKey(tr, k))
        return SymKeyArray(out)

    def bits(self, shape=()):
        return self._shaped(shape, lambda per, j: Sym(
            self.tr, self.tr.tmp("int32_t", f"(int32_t){self._bits(per, j)}"), "i32", True))

    def uniform(self, minval, maxval, shape=()):
        tr = self.tr
        if isinstance(minval, (Sym, SymVec, torch.Tensor)) or \
                isinstance(maxval, (Sym, SymVec, torch.Tensor)):
            f = self.uniform(0.0, 1.0, shape)
            lo, hi = tr.bound(minval), tr.bound(maxval)
            return tr.torch_call(torch.maximum, (lo, f * (hi - lo) + lo), {})
        lo = np.float32(float(minval))
        span = np.float32(np.float32(float(maxval)) - lo)

        def one(per, j):
            expr = (f"abm::scale_uniform(abm::unit_float({self._bits(per, j)}), "
                    f"{_lit(float(lo), 'f32')}, {_lit(float(span), 'f32')})")
            return Sym(tr, tr.tmp("float", expr), "f32", True)

        return self._shaped(shape, one)

    def normal(self, shape=()):
        return self._shaped(shape, lambda per, j: Sym(
            self.tr, self.tr.tmp("float", f"abm::normal_from_bits({self._bits(per, j)})"),
            "f32", True))

    def randint(self, lo: int, hi: int, shape=()):
        tr = self.tr
        span = 1 if hi <= lo else hi - lo
        a, b = tr.fresh("k"), tr.fresh("k")
        tr.lines.append(f"abm::Key {a}, {b}; abm::split2({self.name}, {a}, {b}, {tr.mode});")

        def one(per, j):
            hb = tr.tmp("uint32_t", f"abm::bits_at({a}, {per}LL, {j}LL, {tr.mode})")
            lb = tr.tmp("uint32_t", f"abm::bits_at({b}, {per}LL, {j}LL, {tr.mode})")
            expr = f"(int32_t)((uint32_t)({lo}) + abm::randint_offset({hb}, {lb}, {span}u))"
            return Sym(tr, tr.tmp("int32_t", expr), "i32", True)

        return self._shaped(shape, one)

    def gamma(self, a, b, kind: int, shape=()):
        """gamma / loggamma (b None) or beta draws: the device sampler the
        library's kernel runs, with float32 parameters per element."""
        from ._lib import BETA, LOGGAMMA

        tr = self.tr
        shape = tuple(shape)
        params = [tr.cast(tr.voperand(v), "f32") for v in ((a,) if b is None else (a, b))]
        for p in params:
            if isinstance(p, SymVec) and (shape == () or len(p) != shape[0]):
                raise Unsupported("a draw parameter that does not match the draw's shape")

        def one(per, j):
            ps = [tr._elem(p, j).name for p in params]
            if kind == BETA:
                expr = f"abm::beta_draw({self.name}, {per}LL, {j}LL, {ps[0]}, {ps[1]}, {tr.mode})"
            else:
                log = "true" if kind == LOGGAMMA else "false"
                expr = f"abm::gamma_draw({self.name}, {per}LL, {j}LL, {ps[0]}, {log}, {tr.mode})"
            return Sym(tr, tr.tmp("float", expr), "f32", True)

        return self._shaped(shape, one)


class SymKeyArray:
    """The keys of ``split(key, num)``."""

    _abm_symkey = True

    def __init__(self, keys: List[SymKey]):
        self.keys = keys

    @property
    def shape(self):
        return torch.Size((len(self.keys), 2))

    def __len__(self):
        return len(self.keys)

    def __iter__(self):
        return iter(self.keys)

    def __getitem__(self, i):
        if isinstance(i, int):
            return self.keys[i]
        raise Unsupported("non-integer index into a key array")


_MAX_VEC = 16   # per-agent vectors up to this length are unrolled


class SymVec:
    """A traced per-agent (or uniform) vector of static length k <= 16: its
    elements as traced scalars of one dtype.  Elementwise operations are
    unrolled; promotion is decided on the whole vector (a per-agent vector
    is a 2-D tensor under vmap, a uniform one 1-D)."""

    _abm_sym = True
    __slots__ = ("tr", "items")

    def __init__(self, tr: "_Trace", items: List[Sym]):
        if not items or len(items) > _MAX_VEC:
            raise Unsupported("a per-agent vector longer than 16")
        self.tr, self.items = tr, list(items)

    weak = False

    @property
    def kind(self) -> str:
        return self.items[0].kind

    @property
    def varying(self) -> bool:
        return any(x.varying for x in self.items)

    @property
    def dtype(self):
        return _TORCH[self.kind]

    @property
    def shape(self):
        return torch.Size((len(self.items),))

    def dim(self):
        return 1

    def size(self, *a):
        return self.shape if not a else len(self.items)

    def __len__(self):
        return len(self.items)

    def __iter__(self):
        return iter(self.items)

    def is_floating_point(self):
        return self.kind in _FLOATS

    def as_array(self):
        return self

    def __getitem__(self, i):
        if isinstance(i, bool) or not isinstance(i, (int, slice)):
            raise Unsupported("indexing a per-agent vector by other than an int or a slice")
        if isinstance(i, slice):
            return SymVec(self.tr, self.items[i])
        return self.items[i]

    @property
    def at(self):
        return _SymVecAt(self)

    def to(self, *args, **kwargs):
        return SymVec(self.tr, [x.to(*args, **kwargs) for x in self.items])

    def astype(self, dt):
        return SymVec(self.tr, [x.astype(dt) for x in self.items])

    def float(self):
        return self.to(torch.float32)

    def int(self):
        return self.to(torch.int32)

    def long(self):
        return self.to(torch.int64)

    def bool(self):
        return self.to(torch.bool)

    def clone(self):
        return self

    def contiguous(self):
        return self

    def detach(self):
        return self

    def __bool__(self):
        raise Unsupported("Python control flow on a traced value")

    def __float__(self):
        raise Unsupported("float() of a traced value")

    def item(self):
        raise Unsupported("item() of a traced value")

    def tolist(self):
        raise Unsupported("tolist() of a traced value")

    def sum(self, *a, **k):
        if a or k:
            raise Unsupported("a reduction with arguments over a per-agent vector")
        return self.tr.vsum(self)

    def mean(self, *a, **k):
        if a or k:
            raise Unsupported("a reduction with arguments over a per-agent vector")
        return self.tr.vsum(self, mean=True)

    def prod(self, *a, **k):
        if a or k:
            raise Unsupported("a reduction with arguments over a per-agent vector")
        return self.tr.vprod(self)

    def reshape(self, *shape):
        if shape in ((-1,), ((-1,),), (len(self.items),), ((len(self.items),),)):
            return self
        raise Unsupported("reshape of a per-agent vector")

    def __matmul__(self, other):
        return self.tr.vdot(self, other)

    def __rmatmul__(self, other):
        return self.tr.vdot(other, self)

    def max(self, *a, **k):
        if a or k:
            raise Unsupported("a reduction with arguments over a per-agent vector")
        return self.tr.vextreme(self, True, along_dim=False)

    def min(self, *a, **k):
        if a or k:
            raise Unsupported("a reduction with arguments over a per-agent vector")
        return self.tr.vextreme(self, False, along_dim=False)

    __add__ = Sym.__add__
    __radd__ = Sym.__radd__
    __sub__ = Sym.__sub__
    __rsub__ = Sym.__rsub__
    __mul__ = Sym.__mul__
    __rmul__ = Sym.__rmul__
    __truediv__ = Sym.__truediv__
    __rtruediv__ = Sym.__rtruediv__
    __floordiv__ = Sym.__floordiv__
    __rfloordiv__ = Sym.__rfloordiv__
    __mod__ = Sym.__mod__
    __rmod__ = Sym.__rmod__
    __pow__ = Sym.__pow__
    __lt__ = Sym.__lt__
    __le__ = Sym.__le__
    __gt__ = Sym.__gt__
    __ge__ = Sym.__ge__
    __eq__ = Sym.__eq__
    __ne__ = Sym.__ne__
    __and__ = Sym.__and__
    __rand__ = Sym.__rand__
    __or__ = Sym.__or__
    __ror__ = Sym.__ror__
    __xor__ = Sym.__xor__
    __rxor__ = Sym.__rxor__
    __neg__ = Sym.__neg__
    __pos__ = Sym.__pos__
    __abs__ = Sym.__abs__
    __invert__ = Sym.__invert__
    __hash__ = object.__hash__
    __torch_function__ = Sym.__torch_function__


# ---------------------------------------------------------------- the trace
def _unwrap(x):
    if isinstance(x, SymArray):
        return x.t
    if isinstance(x, (list, tuple)):
        return type(x)(_unwrap(v) for v in x)
    if isinstance(x, dict):
        return {k: _unwrap(v) for k, v in x.items()}
    return x


class SymArray:
    """A collection's whole state array, ``model_state['agents_<name>'][var]``
    of shape (M,) or (M, k), read inside a per-agent function
    (jaxabm/model.py:101-116 puts every collection's states there).

    * Indexing it by a traced (per-agent) integer is a gather with JAX's
      meaning: a negative index counts from the end, and an index still out of
      range is clamped to the nearest end.
    * Anything else -- a constant index, ``jnp.mean(arr)``, ``arr > 0.5`` --
      is the same for every agent, as under vmap where the array is not
      batched: it runs as the torch op on the real array, before the kernel,
      and its result enters the kernel as a uniform value."""

    _abm_sym = True
    weak = False
    varying = False

    def __init__(self, tr: "_Trace", t: torch.Tensor):
        self.tr, self.t = tr, t

    def as_array(self):
        return self.t

    def __getitem__(self, i):
        first = i[0] if isinstance(i, tuple) and i else i
        if not isinstance(first, (Sym, SymVec)):
            return self.t[_unwrap(i)]
        col = None
        if isinstance(i, tuple):
            if len(i) != 2 or self.t.dim() != 2 or isinstance(i[1], bool) or \
                    not isinstance(i[1], int):
                raise Unsupported("an agents-array index other than [j] or [j, constant]")
            i, col = i
        if isinstance(i, SymVec):
            return SymVec(self.tr, [self.tr.gather(self.t, v, col) for v in i.items])
        return self.tr.gather(self.t, i, col)

    def __getattr__(self, name):
        return getattr(self.t, name)

    def __len__(self):
        return len(self.t)

    def __iter__(self):
        return iter(self.t)

    @classmethod
    def __torch_function__(cls, func, types, args=(), kwargs=None):
        return func(*_unwrap(args), **_unwrap(kwargs or {}))


def _delegate(name: str):
    def op(self, *a):
        return getattr(self.t, name)(*_unwrap(a))
    return op


for _name in ("__add__", "__radd__", "__sub__", "__rsub__", "__mul__", "__rmul__",
              "__truediv__", "__rtruediv__", "__floordiv__", "__rfloordiv__", "__mod__",
              "__rmod__", "__pow__", "__rpow__", "__lt__", "__le__", "__gt__", "__ge__",
              "__eq__", "__ne__", "__and__", "__rand__", "__or__", "__ror__", "__xor__",
              "__rxor__", "__neg__", "__abs__", "__invert__", "__bool__", "__float__",
              "__int__", "__index__"):
    setattr(SymArray, _name, _delegate(_name))
SymArray.__hash__ = object.__hash__


class _SymVecAt:
    """``v.at[j]`` on a per-agent vector with a constant index: ``set``,
    ``add``, ``multiply`` and ``get``, as torch's index updates compute them
    (the value is converted to the vector's dtype first)."""

    def __init__(self, v: "SymVec", j=None):
        self.v, self.j = v, j

    def __getitem__(self, j):
        if isinstance(j, bool) or not isinstance(j, int) or not -len(self.v) <= j < len(self.v):
            raise Unsupported("a vector update at other than a constant in-range index")
        return _SymVecAt(self.v, j % len(self.v))

    def _update(self, fn):
        items = list(self.v.items)
        items[self.j] = fn(items[self.j])
        return SymVec(self.v.tr, items)

    def get(self):
        return self.v.items[self.j]

    def set(self, value):
        tr = self.v.tr
        return self._update(lambda e: tr.cast(tr.operand(value), e.kind))

    def add(self, value):
        tr = self.v.tr
        return self._update(lambda e: tr.binary("add", e, tr.cast(tr.operand(value), e.kind)))

    def multiply(self, value):
        tr = self.v.tr
        return self._update(lambda e: tr.binary("mul", e, tr.cast(tr.operand(value), e.kind)))

    mul = multiply


class _Trace:
    def __init__(self, n: int, mode: int, allow_host: bool = False):
        self.n = n
        self.mode = int(mode)
        # host tensors accepted as env values (tracing for a compile check only)
        self.allow_host = allow_host
        self.lines: List[str] = []
        self.words: List[Tuple[str, Any]] = []   # ("ptr", tensor) | ("f64", v) | ("i64", v)
        self.count = 0

    def fresh(self, prefix: str = "v") -> str:
        self.count += 1
        return f"{prefix}{self.count}"

    def tmp(self, ct: str, expr: str) -> str:
        name = self.fresh()
        self.lines.append(f"const {ct} {name} = {expr};")
        return name

    def word(self, kind: str, value) -> int:
        self.words.append((kind, value))
        return len(self.words) - 1

    # -- leaves
    def agent_input(self, t: torch.Tensor):
        kind = _KIND.get(t.dtype)
        if (kind is None or t.dim() not in (1, 2) or t.shape[0] != self.n or
                not t.is_contiguous() or (t.dim() == 2 and not 0 < t.shape[1] <= _MAX_VEC)):
            raise Unsupported("a state variable that is not a scalar or short vector per agent")
        w = self.word("ptr", t)
        if t.dim() == 1:
            name = self.tmp(_CT[kind], f"((const {_CT[kind]} *)a.w[{w}])[i]")
            return Sym(self, name, kind, True)
        k = t.shape[1]
        return SymVec(self, [Sym(self, self.tmp(_CT[kind], f"((const {_CT[kind]} *)a.w[{w}])"
                                                        f"[i * {k}LL + {j}]"), kind, True)
                             for j in range(k)])

    def uniform_tensor(self, t: torch.Tensor) -> Sym:
        kind = _KIND.get(t.dtype)
        if kind is None or t.dim() != 0 or not (t.is_cuda or self.allow_host):
            raise Unsupported("an env tensor that is not a 0-d device scalar")
        w = self.word("ptr", t)
        return Sym(self, self.tmp(_CT[kind], f"*((const {_CT[kind]} *)a.w[{w}])"), kind, False)

    def python_param(self, v) -> Sym:
        if isinstance(v, bool):
            w = self.word("i64", int(v))
            return Sym(self, self.tmp("bool", f"(a.w[{w}] != 0ull)"), "pyb", False)
        if isinstance(v, int):
            if not -(1 << 63) <= v < (1 << 63):
                raise Unsupported("Python int out of int64 range")
            w = self.word("i64", v)
            return Sym(self, self.tmp("int64_t", f"(int64_t)a.w[{w}]"), "pyi", False)
        w = self.word("f64", float(v))
        return Sym(self, self.tmp("double", f"__longlong_as_double((long long)a.w[{w}])"),
                   "pyf", False)

    def const(self, v) -> Sym:
        if isinstance(v, bool):
            return Sym(self, _lit(v, "pyb"), "pyb", False)
        if isinstance(v, int):
            return Sym(self, _lit(v, "pyi"), "pyi", False)
        if isinstance(v, float):
            return Sym(self, _lit(v, "pyf"), "pyf", False)
        raise Unsupported(f"constant of type {type(v).__name__}")

    def operand(self, x) -> Sym:
        if isinstance(x, Sym):
            if x.tr is not self:
                raise Unsupported("values of two traces")
            return x
        if isinstance(x, (bool, int, float)):
            return self.const(x)
        if isinstance(x, np.generic) and not isinstance(x, np.ndarray):
            kind = _KIND.get(torch.from_numpy(np.asarray(x)).dtype)
            if kind is None:
                raise Unsupported("numpy scalar type")
            s = self.const(x.item())
            return self.cast(s, kind) if not s.weak or kind not in ("f64", "i64") else s
        if isinstance(x, torch.Tensor) and x.dim() == 0 and (x.is_cuda or self.allow_host):
            return self.uniform_tensor(x)
        raise Unsupported(f"operand of type {type(x).__name__}")

    # -- vectors
    def voperand(self, x):
        """A traced scalar or vector operand (vectors: SymVec or a small 1-D
        device tensor, read as a uniform vector)."""
        if isinstance(x, SymVec):
            if x.tr is not self:
                raise Unsupported("values of two traces")
            return x
        if (isinstance(x, torch.Tensor) and x.dim() == 1 and (x.is_cuda or self.allow_host) and
                0 < x.shape[0] <= _MAX_VEC):
            return self.uniform_vector(x)
        return self.operand(x)

    def uniform_vector(self, t: torch.Tensor) -> SymVec:
        kind = _KIND.get(t.dtype)
        if kind is None or not t.is_contiguous():
            raise Unsupported("an env vector of an unsupported dtype")
        w = self.word("ptr", t)
        return SymVec(self, [Sym(self, self.tmp(_CT[kind], f"((const {_CT[kind]} *)a.w[{w}])[{j}]"),
                                 kind, False) for j in range(t.shape[0])])

    @staticmethod
    def _vstand_in(x):
        if isinstance(x, SymVec):
            return torch.empty((1, len(x)) if x.varying else (len(x),), dtype=_TORCH[x.kind])
        return _Trace._stand_in(x)

    def _vlen(self, *xs) -> int:
        k = 1
        for x in xs:
            if isinstance(x, SymVec):
                if k not in (1, len(x)):
                    raise Unsupported("vectors of different lengths")
                k = len(x)
        return k

    @staticmethod
    def _elem(x, j):
        return x.items[j] if isinstance(x, SymVec) else x

    def _vkind(self, a, b) -> str:
        if not isinstance(a, SymVec) and not isinstance(b, SymVec):
            return self.promote(a, b)
        return _KIND[torch.result_type(self._vstand_in(a), self._vstand_in(b))]

    def _vcast(self, x, kind):
        """Cast a non-weak operand to the vector's promoted type (a Python
        scalar stays one: torch converts it inside the operation)."""
        return x if x.weak or kind in _WEAK else self.cast(x, kind)

    def vmap2(self, fn, a, b):
        a, b = self.voperand(a), self.voperand(b)
        if not isinstance(a, SymVec) and not isinstance(b, SymVec):
            return fn(a, b)
        kind = self._vkind(a, b)
        k = self._vlen(a, b)
        return SymVec(self, [fn(self._vcast(self._elem(a, j), kind),
                                self._vcast(self._elem(b, j), kind)) for j in range(k)])

    @staticmethod
    def _sum_slots(k: int) -> int:
        """Accumulators torch's reduction uses for a row of k contiguous
        float32 values: the largest power of two <= k, at most 16."""
        a = 1
        while a * 2 <= k and a < 16:
            a *= 2
        return a

    def _f32_sum(self, names: list[str]) -> str:
        """The float32 sum of the values `names` in the order torch's
        reduction adds a row of k = len(names) <= 16 contiguous float32
        values (measured with torch 2.10 on ROCm for k = 1..20, also under
        vmap): A = _sum_slots(k) accumulators start from zero, value j is
        added to accumulator j mod A, and the accumulators are combined in
        adjacent pairs, ((a0 + a1) + (a2 + a3)) + ...  For k <= 7 that is
        (0 + x0) + x1, ((0 + x0) + x2) + (0 + x1), and four accumulators."""
        A = self._sum_slots(len(names))
        acc = ["0.0f"] * A
        for j, v in enumerate(names):
            acc[j % A] = self.tmp("float", f"({acc[j % A]} + {v})")
        while len(acc) > 1:
            acc = [self.tmp("float", f"({acc[i]} + {acc[i + 1]})") for i in range(0, len(acc), 2)]
        return acc[0]

    def vsum(self, x: SymVec, mean: bool = False, out_kind: Optional[str] = None) -> Sym:
        """torch.sum / torch.mean over a per-agent vector: float32 values in
        torch's order (_f32_sum); a mean multiplies the sum by float32(1/k).
        Integer and boolean sums are int64, or int32 (wrapping) when asked
        for: exact, so any order."""
        k, kind = len(x), x.kind
        if kind in ("bool", "u8", "i32", "i64"):
            if mean:
                raise Unsupported("mean of an integer vector")
            acc = self.cast(x.items[0], "i64")
            for v in x.items[1:]:
                acc = Sym(self, self.tmp("int64_t", f"({acc.name} + (int64_t)({v.name}))"), "i64",
                          x.varying)
            if out_kind == "i32":
                acc = Sym(self, self.tmp("int32_t", f"(int32_t)(uint32_t)(uint64_t){acc.name}"),
                          "i32", x.varying)
            return acc
        if kind != "f32" or not 1 <= k <= _MAX_VEC:
            raise Unsupported("a reduction over a vector other than 1..16 float32 values")
        t = self._f32_sum([v.name for v in x.items])
        if mean:
            t = self.tmp("float", f"({t} * {_lit(float(np.float32(1.0 / k)), 'f32')})")
        return Sym(self, t, "f32", x.varying)

    def vprod(self, x: SymVec, out_kind: Optional[str] = None) -> Sym:
        """torch.prod over a per-agent vector.  float32: the order torch's
        reduction multiplies a row of k <= 16 values in (measured with torch
        2.10 on ROCm for k = 1..20, also under vmap): 16 accumulators
        starting from one, value j in accumulator j, combined by halves,
        a[i] * a[i + 8], then a[i] * a[i + 4], a[i] * a[i + 2], a[0] * a[1]
        (products with an accumulator still at one are exact and left
        out).  Integers and booleans: int64 (or int32 when asked for),
        wrapping, exact in any order."""
        k, kind = len(x), x.kind
        if kind in ("bool", "u8", "i32", "i64"):
            acc = f"(uint64_t)(int64_t)({x.items[0].name})"
            for v in x.items[1:]:
                acc = self.tmp("uint64_t", f"({acc} * (uint64_t)(int64_t)({v.name}))")
            if out_kind == "i32":
                return Sym(self, self.tmp("int32_t", f"(int32_t)(uint32_t){acc}"), "i32",
                           x.varying)
            return Sym(self, self.tmp("int64_t", f"(int64_t){acc}"), "i64", x.varying)
        if kind != "f32" or not 1 <= k <= _MAX_VEC:
            raise Unsupported("a product over a vector other than 1..16 float32 values")
        slots: list = [v.name for v in x.items] + [None] * (16 - k)
        while len(slots) > 1:
            h = len(slots) // 2
            nxt = []
            for i in range(h):
                lo, hi = slots[i], slots[i + h]
                if lo is None or hi is None:
                    nxt.append(hi if lo is None else lo)
                else:
                    nxt.append(self.tmp("float", f"({lo} * {hi})"))
            slots = nxt
        return Sym(self, slots[0], "f32", x.varying)

    def vnorm(self, x: SymVec) -> Sym:
        """torch.linalg.norm / vector_norm (2-norm) of a per-agent float32
        vector: sqrt of the squares summed in torch's order (each square
        rounded, then added: no fused multiply-add), as measured for
        k = 1..20."""
        if x.kind != "f32" or not 1 <= len(x) <= _MAX_VEC:
            raise Unsupported("a norm of a vector other than 1..16 float32 values")
        sq = [self.tmp("float", f"({v.name} * {v.name})") for v in x.items]
        return Sym(self, self.tmp("float", f"sqrtf({self._f32_sum(sq)})"), "f32", x.varying)

    def vvar(self, x: SymVec, correction: float, take_sqrt: bool) -> Sym:
        """torch.var / torch.std over a per-agent float32 vector, as torch's
        Welford reduction gives them (measured with torch 2.10 on ROCm for
        k = 1..20, also under vmap).  A = _sum_slots(k) lanes; lane l holds
        value l and, when l + A < k, value l + A, each started from the
        empty accumulator and the two merged; the lanes merge in adjacent
        pairs.  Merges are Chan et al.'s pairwise update with its two
        multiply-adds fused (abm_wf_merge); the result is m2 / (n - c), or
        m2 / 0 when n <= c, and its square root for std."""
        k = len(x)
        if x.kind != "f32" or not 1 <= k <= _MAX_VEC:
            raise Unsupported("a variance of a vector other than 1..16 float32 values")
        A = self._sum_slots(k)
        names = [v.name for v in x.items]
        lanes = []
        for lane in range(A):
            acc = self.tmp("abm_welford", f"abm_wf_first({names[lane]})")
            if lane + A < k:
                acc = self.tmp("abm_welford",
                               f"abm_wf_merge({acc}, abm_wf_first({names[lane + A]}))")
            lanes.append(acc)
        while len(lanes) > 1:
            lanes = [self.tmp("abm_welford", f"abm_wf_merge({lanes[i]}, {lanes[i + 1]})")
                     for i in range(0, len(lanes), 2)]
        c = _lit(float(np.float32(correction)), "f32")
        v = self.tmp("float", f"abm_wf_var({lanes[0]}, {c})")
        if take_sqrt:
            v = self.tmp("float", f"sqrtf({v})")
        return Sym(self, v, "f32", x.varying)

    def vdot(self, a, b) -> Sym:
        """torch.dot / torch.matmul / ``@`` of two per-agent float32 vectors
        of 1..15 values, as torch's batched product gives them (measured with
        torch 2.10 on ROCm for k = 1..20 at 1 to 262,144 agents, with signed
        zeros): one fused multiply-add chain from +0, value j after value
        j - 1.  From 16 values on torch's product takes another order, so
        those, and products with an env vector (another batching), are not
        traced."""
        a, b = self.voperand(a), self.voperand(b)
        if not (isinstance(a, SymVec) and isinstance(b, SymVec) and a.varying and b.varying and
                a.kind == "f32" and b.kind == "f32" and len(a) == len(b) and
                1 <= len(a) <= 15):
            raise Unsupported("a dot product other than of two per-agent float32 vectors of "
                              "1..15 values")
        acc = "0.0f"
        for x, y in zip(a.items, b.items):
            acc = self.tmp("float", f"__builtin_fmaf({x.name}, {y.name}, {acc})")
        return Sym(self, acc, "f32", True)

    def vcumsum(self, x: SymVec, out_kind: Optional[str] = None) -> SymVec:
        """torch.cumsum over a per-agent vector.  float32: torch's scan of a
        row (measured with torch 2.10 on ROCm for k = 1..20, also under
        vmap): the first value plus the zero carry-in, then Sklansky levels
        m = 0, 1, ...: with s = 2^m, every position in the upper half of a
        block of 2s adds the last position of the lower half.  Integers and
        booleans: int64 (int32, wrapping, when asked for), exact in any
        order."""
        k, kind = len(x), x.kind
        if kind in ("bool", "u8", "i32", "i64"):
            ok = out_kind or "i64"
            ct = _CT[ok]
            acc, out = None, []
            for v in x.items:
                term = f"({ct})({v.name})"
                acc = self.tmp(ct, term if acc is None else
                               (f"({ct})((uint64_t)(int64_t){acc} + (uint64_t)(int64_t){term})"
                                if ok == "i64" else
                                f"({ct})((uint32_t){acc} + (uint32_t){term})"))
                out.append(Sym(self, acc, ok, x.varying))
            return SymVec(self, out)
        if kind != "f32" or not 1 <= k <= _MAX_VEC:
            raise Unsupported("a cumulative sum of a vector other than 1..16 float32 values")
        buf = [v.name for v in x.items]
        buf[0] = self.tmp("float", f"({buf[0]} + 0.0f)")
        m = 0
        while (1 << m) < k:
            s = 1 << m
            for t in range(k):
                a = ((t >> m) << (m + 1)) | s
                ti, si = a + (t % s), a - 1
                if ti < k:
                    buf[ti] = self.tmp("float", f"({buf[ti]} + {buf[si]})")
            m += 1
        return SymVec(self, [Sym(self, b, "f32", x.varying) for b in buf])

    def vsort(self, x: SymVec) -> tuple:
        """torch.sort(stable=True) / torch.argsort(stable=True) of a
        per-agent vector: (values, int64 indices).  float32 keys are ordered
        as torch's radix sort orders them (measured with torch 2.10 on ROCm
        for k = 1..16 on rows of ties, signed zeros, infinities and NaNs of
        both signs, also under vmap): -0 counts as +0, then the bits are
        mapped to an unsigned order (abm_sort_key), so a NaN with its sign
        bit set comes first and one without it last; equal keys keep their
        order.  Integers and booleans: by value, stable.  The rank of value
        i is the number of values placed before it; output p selects the
        value of rank p."""
        k, kind = len(x), x.kind
        if kind not in ("bool", "u8", "i32", "i64", "f32"):
            raise Unsupported("a sort of this dtype")
        ct = _CT[kind]
        if kind == "f32":
            keys = [self.tmp("uint32_t", f"abm_sort_key({v.name})") for v in x.items]
        else:
            keys = [v.name for v in x.items]
        rank = ["0"] * k
        for i in range(k):
            for j in range(i + 1, k):
                after = self.tmp("int", f"(int)({keys[j]} < {keys[i]})")   # j placed before i
                rank[i] = self.tmp("int", f"({rank[i]} + {after})")
                rank[j] = self.tmp("int", f"({rank[j]} + 1 - {after})")
        vals, idxs = [], []
        for p in range(k):
            v, ix = x.items[0].name, "(int64_t)0"
            for i in range(1, k):
                hit = self.tmp("bool", f"({rank[i]} == {p})")
                v = self.tmp(ct, f"({hit} ? {x.items[i].name} : {v})")
                ix = self.tmp("int64_t", f"({hit} ? (int64_t){i} : {ix})")
            vals.append(Sym(self, v, kind, x.varying))
            idxs.append(Sym(self, ix, "i64", x.varying))
        return SymVec(self, vals), SymVec(self, idxs)

    def vany(self, x: SymVec, every: bool) -> Sym:
        """torch.any / torch.all of a per-agent vector: a value counts as
        true when it is nonzero (NaN included)."""
        op = " && " if every else " || "
        terms = [f"((bool)({v.name}))" for v in x.items]
        return Sym(self, self.tmp("bool", "(" + op.join(terms) + ")"), "bool", x.varying)

    def vcount(self, x: SymVec) -> Sym:
        """torch.count_nonzero of a per-agent vector (int64; -0 counts as
        zero, NaN as nonzero)."""
        terms = [f"(int64_t)({v.name} != 0)" for v in x.items]
        return Sym(self, self.tmp("int64_t", "(" + " + ".join(terms) + ")"), "i64", x.varying)

    def vcat(self, parts) -> SymVec:
        """torch.cat / jnp.concatenate of per-agent (or uniform) vectors of
        one dtype, at most 16 values in all."""
        vecs = [self.voperand(p) for p in parts]
        if not vecs or not all(isinstance(v, SymVec) for v in vecs):
            raise Unsupported("a concatenation of other than per-agent vectors")
        if len({v.kind for v in vecs}) != 1:
            raise Unsupported("a concatenation of vectors of different dtypes")
        return SymVec(self, [e for v in vecs for e in v.items])

    def vextreme(self, x: SymVec, largest: bool, along_dim: bool) -> Sym:
        """torch.max / torch.min (whole vector) and torch.amax / torch.amin
        (over its dim) of a per-agent vector, as torch's reductions give
        them (measured with torch 2.10 on ROCm for k = 1..16 with signed
        zeros, infinities and NaN, also under vmap).  A NaN propagates.
        max/min scan the values in order and keep the running value on a
        tie (the first of +0 and -0 wins); amax/amin keep _sum_slots(k)
        running values (value j in slot j mod A), combined in adjacent
        pairs, and take the later value on a tie.  Integers and booleans:
        exact."""
        kind = x.kind
        if kind not in ("bool", "u8", "i32", "i64", "f32", "f64"):
            raise Unsupported("a max/min of this dtype")
        ct = _CT[kind]
        lt = "<" if largest else ">"
        gt = ">" if largest else "<"
        floats = kind in ("f32", "f64")

        def keep_first(a: str, b: str) -> str:   # running a, next value b
            if floats:
                return self.tmp(ct, f"(({a} != {a}) ? {a} : (({b} != {b}) ? {b} : "
                                    f"(({a} {lt} {b}) ? {b} : {a})))")
            return self.tmp(ct, f"(({a} {lt} {b}) ? {b} : {a})")

        def take_later(a: str, b: str) -> str:
            if floats:
                return self.tmp(ct, f"((({a} != {a}) || ({a} {gt} {b})) ? {a} : {b})")
            return self.tmp(ct, f"(({a} {gt} {b}) ? {a} : {b})")

        names = [v.name for v in x.items]
        if not along_dim:
            acc = names[0]
            for v in names[1:]:
                acc = keep_first(acc, v)
            return Sym(self, acc, kind, x.varying)
        A = self._sum_slots(len(names))
        slots: list = [None] * A
        for j, v in enumerate(names):
            slots[j % A] = v if slots[j % A] is None else take_later(slots[j % A], v)
        while len(slots) > 1:
            nxt = []
            for i in range(0, len(slots), 2):
                a, b = slots[i], slots[i + 1]
                nxt.append(b if a is None else (a if b is None else take_later(a, b)))
            slots = nxt
        return Sym(self, slots[0], kind, x.varying)

    def varg(self, x: SymVec, largest: bool) -> Sym:
        """torch.argmax / argmin of a per-agent vector (int64): the first
        index of the largest (smallest) value; a NaN counts as larger than
        everything, so the first NaN's index when there is one."""
        kind = x.kind
        if kind not in ("bool", "u8", "i32", "i64", "f32", "f64"):
            raise Unsupported("an argmax/argmin of this dtype")
        ct = _CT[kind]
        cmp = ">" if largest else "<"
        best, idx = x.items[0].name, "(int64_t)0"
        for j, v in enumerate(x.items[1:], start=1):
            if kind in ("f32", "f64"):
                take = self.tmp("bool", f"(({best} == {best}) && (({v.name} != {v.name}) || "
                                        f"({v.name} {cmp} {best})))")
            else:
                take = self.tmp("bool", f"({v.name} {cmp} {best})")
            best = self.tmp(ct, f"({take} ? {v.name} : {best})")
            idx = self.tmp("int64_t", f"({take} ? (int64_t){j} : {idx})")
        return Sym(self, idx, "i64", x.varying)

    def stack(self, values) -> SymVec:
        """jnp.array / jnp.stack of per-agent or uniform scalars (Python
        scalars become 0-d arrays of JAX's default dtype first)."""
        xs = [self.operand(v).as_array() for v in values]
        if not xs:
            raise Unsupported("an empty vector")
        kind = xs[0].kind
        for x in xs[1:]:
            kind = self._stack_kind(kind, x)
        return SymVec(self, [self.cast(x, kind) for x in xs])

    def _stack_kind(self, kind: str, x: Sym) -> str:
        return _KIND[torch.result_type(torch.empty((1,), dtype=_TORCH[kind]),
                                       torch.empty((1,), dtype=_TORCH[x.kind]))]

    def bound(self, v):
        """random.uniform's tensor-bound branch: bounds as float32 values."""
        if isinstance(v, (Sym, SymVec)):
            return self.cast(v, "f32")
        if isinstance(v, torch.Tensor):
            return self.cast(self.voperand(v), "f32")
        return self.cast(self.const(float(v)), "f32")

    # -- casts and promotion
    def cast(self, x, kind: str):
        if isinstance(x, SymVec):
            return SymVec(self, [self.cast(v, kind) for v in x.items])
        x = self.operand(x)
        if x.kind == kind:
            return x
        if kind in _WEAK:
            raise Unsupported("cast to a Python scalar")
        return Sym(self, self.tmp(_CT[kind], f"(({_CT[kind]})({x.name}))"), kind, x.varying)

    @staticmethod
    def _stand_in(x: Sym):
        if x.weak:
            return {"pyf": 0.0, "pyi": 0, "pyb": False}[x.kind]
        return torch.empty((1,) if x.varying else (), dtype=_TORCH[x.kind])

    def promote(self, a: Sym, b: Sym) -> str:
        if a.weak and b.weak:
            if "pyf" in (a.kind, b.kind):
                return "pyf"
            if "pyi" in (a.kind, b.kind):
                return "pyi"
            return "pyb"
        # torch's promotion depends on the kinds and on which operands are
        # 0-d (uniform) or 1-d (varying) only: memoised
        k = (a.kind, a.varying, b.kind, b.varying)
        r = _PROMOTED.get(k)
        if r is None:
            r = _PROMOTED[k] = _KIND[torch.result_type(self._stand_in(a), self._stand_in(b))]
        return r

    # -- element operations
    def binary(self, op: str, a, b):
        if isinstance(a, (SymVec, torch.Tensor)) or isinstance(b, (SymVec, torch.Tensor)):
            va, vb = self.voperand(a), self.voperand(b)
            if isinstance(va, SymVec) or isinstance(vb, SymVec):
                return self.vmap2(lambda x, y: self.binary(op, x, y), va, vb)
        a, b = self.operand(a), self.operand(b)
        varying = a.varying or b.varying
        kind = self.promote(a, b)
        weak = kind in _WEAK
        if op == "div":
            if kind not in _FLOATS:
                kind = "pyf" if weak else "f32"
        if op in ("lt", "le", "gt", "ge", "eq", "ne"):
            sym = {"lt": "<", "le": "<=", "gt": ">", "ge": ">=", "eq": "==", "ne": "!="}[op]
            if kind == "pyb" and op in ("lt", "le", "gt", "ge"):
                kind = "pyi"
            ca, cb = self._as(a, kind), self._as(b, kind)
            return Sym(self, self.tmp("bool", f"({ca} {sym} {cb})"), "pyb" if weak else "bool",
                       varying)
        if op in ("and", "or", "xor"):
            if kind in ("bool", "pyb"):
                expr = {"and": "{} && {}", "or": "{} || {}", "xor": "{} != {}"}[op]
            elif kind in _FLOATS:
                raise Unsupported("bitwise operation on floats")
            else:
                expr = {"and": "{} & {}", "or": "{} | {}", "xor": "{} ^ {}"}[op]
            ca, cb = self._as(a, kind), self._as(b, kind)
            return Sym(self, self.tmp(_CT[kind], "(" + expr.format(ca, cb) + ")"), kind, varying)
        if op in ("atan2", "hypot"):
            # torch's element functions: atan2f / hypotf of the promoted
            # floats (integers and booleans promote to float32)
            if weak:
                raise Unsupported(f"{op} of two Python scalars")
            if kind not in _FLOATS:
                kind = "f32"
            fn = op + ("f" if kind == "f32" else "")
            ca, cb = self._as(a, kind), self._as(b, kind)
            return Sym(self, self.tmp(_CT[kind], f"{fn}({ca}, {cb})"), kind, varying)
        if kind in ("bool", "pyb"):
            if weak and op in ("add", "sub", "mul", "floordiv", "mod"):
                kind = "pyi"          # Python: True + True == 2
            else:
                raise Unsupported(f"{op} on booleans")
        ca, cb = self._as(a, kind), self._as(b, kind)
        ct = _CT[kind]
        # torch divides a tensor by a Python scalar as a multiply by the
        # scalar's reciprocal (computed in the result type), and a Python
        # scalar by a tensor as reciprocal(tensor) * scalar
        scalar_b = b.weak and not a.weak
        scalar_a = a.weak and not b.weak
        if op == "div" and scalar_b:
            expr = f"({ca} * (({ct})1 / {cb}))"
        elif op == "div" and scalar_a:
            expr = f"((({ct})1 / {cb}) * {ca})"
        elif op in ("add", "sub", "mul", "div"):
            expr = f"({ca} {dict(add='+', sub='-', mul='*', div='/')[op]} {cb})"
        elif op == "floordiv":
            if kind in _FLOATS and scalar_b:
                expr = f"abm_floordiv_fs<{ct}>({ca}, {cb})"
            elif kind in _FLOATS:
                expr = f"abm_floordiv_f<{ct}>({ca}, {cb})"
            else:
                expr = f"abm_floordiv_i<{ct}>({ca}, {cb})"
        elif op == "mod":
            if kind == "pyf":
                expr = f"abm_pymod_f({ca}, {cb})"
            elif kind in _FLOATS:
                expr = f"abm_mod_f<{ct}>({ca}, {cb})"
            else:
                expr = f"abm_mod_i<{ct}>({ca}, {cb})"
        else:
            raise Unsupported(op)
        return Sym(self, self.tmp(ct, expr), kind, varying)

    def _as(self, x: Sym, kind: str) -> str:
        if x.kind == kind:
            return x.name
        return f"(({_CT[kind]})({x.name}))"

    def power(self, x, e):
        """x ** e for a Python number e, as torch's pow of a tensor by a
        scalar: 2 and 3 multiply, 0.5 is sqrt, -0.5 torch's rsqrt (float32
        only), -1 and -2 reciprocals, any other exponent of a float pow(x, e)
        with e in x's dtype; integer bases take only 2 and 3."""
        if isinstance(e, bool) or not isinstance(e, (int, float)):
            raise Unsupported("a power with an exponent that is not a Python number")
        if isinstance(x, SymVec):
            return SymVec(self, [self.power(v, e) for v in x.items])
        x = self.operand(x)
        if x.weak:
            raise Unsupported("a power of a Python scalar")
        kind = self.promote(x, self.const(e))
        cx = self._as(x, kind)
        if e in (2, 3):
            expr = f"({cx} * {cx})" if e == 2 else f"({cx} * {cx} * {cx})"
        elif kind not in ("f32", "f64"):
            raise Unsupported("an integer power other than x**2 or x**3")
        else:
            f32 = kind == "f32"
            one = "1.0f" if f32 else "1.0"
            if e == 0.5:
                expr = f"{'sqrtf' if f32 else 'sqrt'}({cx})"
            elif e == -0.5:
                # torch computes this case as its rsqrt: for float32 that is
                # (float)(1 / sqrt((double)x)) on every bit pattern (rsqrtf
                # and 1 / sqrtf differ from it); float64 stays vmapped
                if not f32:
                    raise Unsupported("x ** -0.5 of a float64 value")
                expr = f"((float)(1.0 / sqrt((double){cx})))"
            elif e == -1:
                expr = f"({one} / {cx})"
            elif e == -2:
                expr = f"({one} / ({cx} * {cx}))"
            else:
                expr = f"{'powf' if f32 else 'pow'}({cx}, {_lit(float(e), kind)})"
        return Sym(self, self.tmp(_CT[kind], expr), kind, x.varying)

    _UNARY_F = {"exp": "expf", "log": "logf", "log1p": "log1pf", "expm1": "expm1f",
                "sqrt": "sqrtf", "sin": "sinf", "cos": "cosf", "tan": "tanf", "tanh": "tanhf",
                "atan": "atanf"}

    def unary(self, op: str, x):
        x = self.voperand(x)
        if isinstance(x, SymVec):
            return SymVec(self, [self.unary(op, v) for v in x.items])
        if x.weak:
            if op == "neg":
                if x.kind == "pyb":
                    raise Unsupported("negation of a bool")
                return Sym(self, self.tmp(_CT[x.kind], f"(-{x.name})"), x.kind, x.varying)
            if op == "abs":
                k = "pyi" if x.kind == "pyb" else x.kind
                fn = "fabs" if k == "pyf" else "llabs"
                return Sym(self, self.tmp(_CT[k], f"{fn}({self._as(x, k)})"), k, x.varying)
            x = x.as_array()
        k, ct = x.kind, _CT[x.kind]
        if op == "neg":
            if k == "bool":
                raise Unsupported("negation of a bool tensor")
            return Sym(self, self.tmp(ct, f"(({ct})(-{x.name}))"), k, x.varying)
        if op == "abs":
            if k == "bool":
                raise Unsupported("abs of a bool tensor")
            expr = (f"fabsf({x.name})" if k == "f32" else f"fabs({x.name})" if k == "f64"
                    else f"(({ct})({x.name} < 0 ? -{x.name} : {x.name}))")
            return Sym(self, self.tmp(ct, expr), k, x.varying)
        if op == "invert":
            expr = f"(!{x.name})" if k == "bool" else f"(({ct})(~{x.name}))"
            if k in _FLOATS:
                raise Unsupported("~ on floats")
            return Sym(self, self.tmp(ct, expr), k, x.varying)
        if op == "rsqrt":
            # torch.rsqrt of float32 equals (float)(1 / sqrt((double)x)) on
            # every float32 bit pattern (measured: torch 2.10, ROCm, MI355X);
            # float64 is left to the vmapped path
            if k not in _FLOATS:
                x = self.cast(x, "f32")
            if x.kind != "f32":
                raise Unsupported("rsqrt of a float64 value")
            return Sym(self, self.tmp("float", f"((float)(1.0 / sqrt((double){x.name})))"),
                       "f32", x.varying)
        if op in self._UNARY_F:
            if k not in _FLOATS:
                x = self.cast(x, "f32")
            if x.kind == "f64":
                return Sym(self, self.tmp("double", f"{self._UNARY_F[op][:-1]}({x.name})"),
                           "f64", x.varying)
            return Sym(self, self.tmp("float", f"{self._UNARY_F[op]}({x.name})"), "f32",
                       x.varying)
        if op in ("floor", "ceil", "round", "trunc"):
            if k not in _FLOATS:
                return x
            fn = {"floor": "floor", "ceil": "ceil", "round": "rint", "trunc": "trunc"}[op]
            fn = fn + ("f" if k == "f32" else "")
            return Sym(self, self.tmp(ct, f"{fn}({x.name})"), k, x.varying)
        if op == "square":
            return self.binary("mul", x, x)
        if op in ("isnan", "isfinite"):
            if k not in _FLOATS:
                return Sym(self, "true" if op == "isfinite" else "false", "bool", x.varying)
            fn = "isnan" if op == "isnan" else "isfinite"
            return Sym(self, self.tmp("bool", f"{fn}({x.name})"), "bool", x.varying)
        if op == "sign":
            if k == "bool":
                return x
            expr = f"(({ct})(({x.name} > 0) - ({x.name} < 0)))"   # 0 for NaN, as torch
            return Sym(self, self.tmp(ct, expr), k, x.varying)
        if op == "logical_not":
            return Sym(self, self.tmp("bool", f"(!(bool)({x.name}))"), "bool", x.varying)
        raise Unsupported(op)

    def where(self, c, a, b):
        c, a, b = self.voperand(c), self.voperand(a), self.voperand(b)
        if any(isinstance(v, SymVec) for v in (c, a, b)):
            kind = self._vkind(a, b)
            if kind in _WEAK:
                kind = _DEFAULT[kind]
            k = self._vlen(c, a, b)
            return SymVec(self, [self.where(self._elem(c, j), self._vcast(self._elem(a, j), kind),
                                            self._vcast(self._elem(b, j), kind))
                                 for j in range(k)])
        kind = self.promote(a, b)
        if kind in _WEAK:
            kind = _DEFAULT[kind]
        cb = c.name if c.kind in ("bool", "pyb") else f"((bool)({c.name}))"
        expr = f"({cb} ? {self._as(a, kind)} : {self._as(b, kind)})"
        return Sym(self, self.tmp(_CT[kind], expr), kind, c.varying or a.varying or b.varying)

    def minmax(self, op: str, a, b):
        a, b = self.voperand(a), self.voperand(b)
        if isinstance(a, SymVec) or isinstance(b, SymVec):
            return self.vmap2(lambda x, y: self.minmax(op, x, y), a, b)
        if a.weak and b.weak:
            a = a.as_array()
        kind = self.promote(a, b)
        ca, cb = self._as(a, kind), self._as(b, kind)
        fn = "abm_min" if op == "min" else "abm_max"
        return Sym(self, self.tmp(_CT[kind], f"{fn}<{_CT[kind]}>({ca}, {cb})"), kind,
                   a.varying or b.varying)

    def clamp(self, x, lo, hi):
        x = self.voperand(x)
        if isinstance(x, SymVec):
            for v in (lo, hi):
                if isinstance(v, (SymVec, torch.Tensor)) and getattr(v, "dim", lambda: 0)() > 0:
                    raise Unsupported("clamp with vector bounds")
            return SymVec(self, [self.clamp(v, lo, hi) for v in x.items])
        if x.weak:
            x = x.as_array()
        kind = x.kind
        for v in (lo, hi):
            if v is not None:
                kind = self.promote(Sym(self, "", kind, x.varying), self.operand(v))
        out = self.cast(x, kind)
        if lo is not None:
            out = Sym(self, self.tmp(_CT[kind], f"abm_clamp_lo<{_CT[kind]}>({out.name}, "
                                                f"{self._as(self.operand(lo), kind)})"),
                      kind, x.varying)
        if hi is not None:
            out = Sym(self, self.tmp(_CT[kind], f"abm_clamp_hi<{_CT[kind]}>({out.name}, "
                                                f"{self._as(self.operand(hi), kind)})"),
                      kind, x.varying)
        return out

    def gather(self, t: torch.Tensor, i, col=None):
        """t[i] (or t[i, col]) for an agents array t of shape (M,) or (M, k)."""
        kind = _KIND.get(t.dtype)
        if kind is None or not t.is_contiguous() or t.dim() not in (1, 2) or t.shape[0] == 0 or \
                (t.dim() == 2 and not 0 < t.shape[1] <= _MAX_VEC):
            raise Unsupported("a gather from an array that is not (M,) or (M, k <= 16)")
        j = self.operand(i)
        if j.kind not in ("i32", "i64", "pyi"):
            raise Unsupported("a gather index that is not an int32/int64 value")
        m = t.shape[0]
        w = self.word("ptr", t)
        idx = self.tmp("int64_t", f"abm_gather_index((int64_t)({j.name}), {m}LL)")
        ct = _CT[kind]
        base = f"((const {ct} *)a.w[{w}])"
        if t.dim() == 1:
            return Sym(self, self.tmp(ct, f"{base}[{idx}]"), kind, j.varying)
        k = t.shape[1]

        def elem(c):
            return Sym(self, self.tmp(ct, f"{base}[{idx} * {k}LL + {c}]"), kind, j.varying)

        if col is not None:
            c = col + k if col < 0 else col
            if not 0 <= c < k:
                raise Unsupported("a gather column out of range")
            return elem(c)
        return SymVec(self, [elem(c) for c in range(k)])

    def nan_to_num(self, x, nan, posinf, neginf):
        """torch.nan_to_num / jnp.nan_to_num: NaN, +inf and -inf replaced by
        the Python values given (infinities default to the dtype's largest
        finite magnitude), converted to the dtype; other dtypes unchanged."""
        x = self.voperand(x)
        if isinstance(x, SymVec):
            return SymVec(self, [self.nan_to_num(v, nan, posinf, neginf) for v in x.items])
        if x.weak:
            raise Unsupported("nan_to_num of a Python scalar")
        if x.kind not in ("f32", "f64"):
            return x
        big = float(np.finfo(np.float32 if x.kind == "f32" else np.float64).max)
        vals = []
        for v, default in ((nan, 0.0), (posinf, big), (neginf, -big)):
            if v is None:
                v = default
            if isinstance(v, bool) or not isinstance(v, (int, float)):
                raise Unsupported("nan_to_num with a replacement that is not a Python number")
            vals.append(_lit(float(v), x.kind))
        c = x.name
        expr = (f"({c} != {c} ? {vals[0]} : ({c} == __builtin_inf() ? {vals[1]} : "
                f"({c} == -__builtin_inf() ? {vals[2]} : {c})))")
        return Sym(self, self.tmp(_CT[x.kind], expr), x.kind, x.varying)

    # -- torch functions called with traced values
    def torch_call(self, func, args, kwargs):
        name = getattr(func, "__name__", "")
        unary = {"abs": "abs", "absolute": "abs", "neg": "neg", "negative": "neg",
                 "exp": "exp", "log": "log", "log1p": "log1p", "expm1": "expm1",
                 "sqrt": "sqrt", "rsqrt": "rsqrt", "sin": "sin", "cos": "cos", "tan": "tan",
                 "tanh": "tanh",
                 "atan": "atan", "arctan": "atan", "floor": "floor", "ceil": "ceil",
                 "trunc": "trunc", "square": "square", "isnan": "isnan",
                 "isfinite": "isfinite", "sign": "sign", "logical_not": "logical_not"}
        if name in unary and len(args) == 1 and not kwargs:
            return self.unary(unary[name], args[0])
        if name == "round" and len(args) == 1 and kwargs.get("decimals", 0) == 0:
            return self.unary("round", args[0])
        binary = {"add": "add", "sub": "sub", "subtract": "sub", "mul": "mul",
                  "multiply": "mul", "div": "div", "true_divide": "div", "divide": "div",
                  "floor_divide": "floordiv", "remainder": "mod", "lt": "lt", "less": "lt",
                  "le": "le", "gt": "gt", "greater": "gt", "ge": "ge", "eq": "eq", "ne": "ne",
                  "bitwise_and": "and", "bitwise_or": "or", "bitwise_xor": "xor",
                  "atan2": "atan2", "arctan2": "atan2", "hypot": "hypot"}
        if name in binary and len(args) == 2 and not kwargs:
            return self.binary(binary[name], args[0], args[1])
        dunder = {"__add__": "add", "__sub__": "sub", "__mul__": "mul", "__truediv__": "div",
                  "__div__": "div", "__floordiv__": "floordiv", "__mod__": "mod",
                  "__lt__": "lt", "__le__": "le", "__gt__": "gt", "__ge__": "ge",
                  "__eq__": "eq", "__ne__": "ne", "__and__": "and", "__or__": "or",
                  "__xor__": "xor"}
        if name in dunder and len(args) == 2:
            return self.binary(dunder[name], args[0], args[1])
        if name.startswith("__r") and "__" + name[3:] in dunder and len(args) == 2:
            return self.binary(dunder["__" + name[3:]], args[1], args[0])
        if name == "__pow__" and len(args) == 2:
            return self.power(args[0], args[1])
        if name in ("__neg__", "__abs__", "__invert__") and len(args) == 1:
            return self.unary(name.strip("_").replace("invert", "invert"), args[0])
        if name in ("logical_and", "logical_or", "logical_xor") and len(args) == 2:
            a = self.cast(self.operand(args[0]).as_array(), "bool")
            b = self.cast(self.operand(args[1]).as_array(), "bool")
            return self.binary(name.split("_")[1], a, b)
        if name == "where" and len(args) == 3 and not kwargs:
            return self.where(*args)
        if name in ("minimum", "maximum", "fmin", "fmax") and len(args) == 2 and not kwargs:
            if name in ("fmin", "fmax"):
                raise Unsupported(name)
            return self.minmax("min" if name == "minimum" else "max", *args)
        if name in ("clamp", "clip"):
            x = args[0]
            lo = args[1] if len(args) > 1 else kwargs.get("min")
            hi = args[2] if len(args) > 2 else kwargs.get("max")
            return self.clamp(x, lo, hi)
        if name == "pow" and len(args) == 2 and not kwargs:
            return self.power(args[0], args[1])
        if name == "nan_to_num" and 1 <= len(args) <= 4:
            opts = dict(zip(("nan", "posinf", "neginf"), args[1:]))
            opts.update(kwargs)
            if set(opts) - {"nan", "posinf", "neginf"}:
                raise Unsupported("nan_to_num options")
            return self.nan_to_num(args[0], opts.get("nan", 0.0), opts.get("posinf"),
                                   opts.get("neginf"))
        if name in ("sum", "mean") and len(args) == 1 and isinstance(args[0], SymVec):
            dim = kwargs.get("dim")
            dt = kwargs.get("dtype")
            int_in = args[0].kind in ("bool", "u8", "i32")
            if dim in (None, 0, -1, (0,), (-1,), [0], [-1]) and not kwargs.get("keepdim") and \
                    (dt is None or (dt == torch.int32 and int_in and name == "sum")):
                return self.vsum(args[0], mean=name == "mean",
                                 out_kind="i32" if dt == torch.int32 else None)
        if name in ("max", "min") and len(args) == 1 and not kwargs and \
                isinstance(args[0], SymVec):
            return self.vextreme(args[0], name == "max", along_dim=False)
        if name in ("amax", "amin") and len(args) == 1 and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "keepdim"} and kwargs.get("dim") in (0, -1, (0,), (-1,), [0], [-1]) \
                and not kwargs.get("keepdim"):
            return self.vextreme(args[0], name == "amax", along_dim=True)
        if name in ("argmax", "argmin") and len(args) == 1 and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "keepdim"} and kwargs.get("dim") in (None, 0, -1) and \
                not kwargs.get("keepdim"):
            return self.varg(args[0], name == "argmax")
        if name == "prod" and len(args) == 1 and isinstance(args[0], SymVec):
            dim = kwargs.get("dim")
            dt = kwargs.get("dtype")
            int_in = args[0].kind in ("bool", "u8", "i32")
            if dim in (None, 0, -1) and not kwargs.get("keepdim") and \
                    (dt is None or (dt == torch.int32 and int_in)):
                return self.vprod(args[0], out_kind="i32" if dt == torch.int32 else None)
        if name in ("linalg_norm", "linalg_vector_norm") and len(args) == 1 and \
                isinstance(args[0], SymVec) and \
                set(kwargs) <= {"ord", "dim", "keepdim", "dtype"} and \
                kwargs.get("ord") in (None, 2) and kwargs.get("dim") in (None, 0, -1) and \
                not kwargs.get("keepdim") and kwargs.get("dtype") is None:
            return self.vnorm(args[0])
        if name in ("dot", "matmul") and len(args) == 2 and not kwargs and \
                (isinstance(args[0], SymVec) or isinstance(args[1], SymVec)):
            return self.vdot(args[0], args[1])
        if name in ("__matmul__", "__rmatmul__") and len(args) == 2 and \
                (isinstance(args[0], SymVec) or isinstance(args[1], SymVec)):
            return self.vdot(*(args if name == "__matmul__" else args[::-1]))
        if name == "cumsum" and len(args) in (1, 2) and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "dtype"}:
            dim = args[1] if len(args) == 2 else kwargs.get("dim")
            dt = kwargs.get("dtype")
            int_in = args[0].kind in ("bool", "u8", "i32")
            if dim in (0, -1) and (dt is None or (dt == torch.int32 and int_in)):
                return self.vcumsum(args[0], out_kind="i32" if dt == torch.int32 else None)
        if name in ("sort", "argsort") and len(args) in (1, 2) and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "stable", "descending"} and kwargs.get("stable") is True \
                and not kwargs.get("descending"):
            dim = args[1] if len(args) == 2 else kwargs.get("dim", -1)
            if dim in (0, -1):
                values, indices = self.vsort(args[0])
                return indices if name == "argsort" else torch.return_types.sort((values, indices))
        if name in ("any", "all") and len(args) == 1 and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "keepdim"} and kwargs.get("dim") in (None, 0, -1) and \
                not kwargs.get("keepdim"):
            return self.vany(args[0], every=name == "all")
        if name == "count_nonzero" and len(args) == 1 and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim"} and kwargs.get("dim") in (None, 0, -1):
            return self.vcount(args[0])
        if name in ("cat", "concatenate", "concat") and len(args) >= 1 and \
                isinstance(args[0], (list, tuple)) and \
                kwargs.get("dim", args[1] if len(args) > 1 else 0) in (0, -1):
            return self.vcat(args[0])
        if name == "flip" and len(args) in (1, 2) and isinstance(args[0], SymVec):
            dims = args[1] if len(args) == 2 else kwargs.get("dims")
            if dims in ((0,), [0], (-1,), [-1], 0, -1):
                return SymVec(self, list(reversed(args[0].items)))
        if name == "roll" and len(args) in (2, 3) and isinstance(args[0], SymVec) and \
                isinstance(args[1], int) and not isinstance(args[1], bool):
            dims = args[2] if len(args) == 3 else kwargs.get("dims")
            if dims in (None, 0, -1, (0,), [0], (-1,), [-1]):
                items, k = args[0].items, len(args[0])
                sh = args[1] % k
                return SymVec(self, items[k - sh:] + items[:k - sh])
        if name in ("var", "std") and len(args) == 1 and isinstance(args[0], SymVec) and \
                set(kwargs) <= {"dim", "keepdim", "correction", "unbiased"} and \
                kwargs.get("dim") in (None, 0, -1, (0,), (-1,), [0], [-1]) and \
                not kwargs.get("keepdim") and \
                not ("correction" in kwargs and "unbiased" in kwargs):
            corr = kwargs.get("correction")
            if corr is None:
                corr = 1 if kwargs.get("unbiased", True) else 0
            if isinstance(corr, (bool, int, float)) and not isinstance(corr, bool):
                return self.vvar(args[0], float(corr), take_sqrt=name == "std")
        if name == "stack" and len(args) >= 1 and isinstance(args[0], (list, tuple)) and \
                kwargs.get("dim", args[1] if len(args) > 1 else 0) in (0, -1):
            return self.stack(args[0])
        raise Unsupported(f"torch.{name}")


_HELPERS = r"""
// element functions with the semantics of torch's (remainder, floor_divide,
// clamp, minimum, maximum) and of Python's // and % on scalars
template <typename T> __device__ __forceinline__ T abm_mod_i(T a, T b) {
  T r = a % b;
  if (r != 0 && ((r < 0) != (b < 0))) r += b;
  return r;
}
template <typename T> __device__ __forceinline__ T abm_floordiv_i(T a, T b) {
  T q = a / b;
  if ((a % b != 0) && ((a < 0) != (b < 0))) q -= 1;
  return q;
}
template <typename T> __device__ __forceinline__ T abm_mod_f(T a, T b) {
  T m = fmod(a, b);
  if (m != 0 && ((b < 0) != (m < 0))) m += b;
  return m;
}
__device__ __forceinline__ double abm_pymod_f(double a, double b) {
  double m = fmod(a, b);
  if (m != 0) {
    if ((b < 0) != (m < 0)) m += b;
  } else {
    m = copysign(0.0, b);
  }
  return m;
}
template <typename T> __device__ __forceinline__ T abm_floordiv_f(T a, T b) {
  if (b == 0) return a / b;
  T m = fmod(a, b);
  T d = (a - m) / b;
  if (m != 0 && ((b < 0) != (m < 0))) d -= 1;
  T f;
  if (d != 0) {
    f = floor(d);
    if (d - f > (T)0.5) f += 1;
  } else {
    f = copysign((T)0, a / b);
  }
  return f;
}
// floor division by a Python scalar: torch multiplies by the reciprocal
template <typename T> __device__ __forceinline__ T abm_floordiv_fs(T a, T b) {
  if (b == 0) return a / b;
  const T inv_b = (T)1 / b;
  T m = fmod(a, b);
  T d = (a - m) * inv_b;
  if (m != 0 && ((b < 0) != (m < 0))) d -= 1;
  T f;
  if (d != 0) {
    f = floor(d);
    if (d - f > (T)0.5) f += 1;
  } else {
    f = copysign((T)0, a * inv_b);
  }
  return f;
}
// a gather index as JAX reads it: negative counts from the end, then clamped
__device__ __forceinline__ int64_t abm_gather_index(int64_t j, int64_t m) {
  if (j < 0) j += m;
  return j < 0 ? 0 : (j >= m ? m - 1 : j);
}
template <typename T> __device__ __forceinline__ bool abm_isnan(T a) { return a != a; }
template <typename T> __device__ __forceinline__ T abm_lo2(T a, T b) { return b < a ? b : a; }
template <typename T> __device__ __forceinline__ T abm_hi2(T a, T b) { return a < b ? b : a; }
template <> __device__ __forceinline__ float abm_lo2<float>(float a, float b) { return fminf(a, b); }
template <> __device__ __forceinline__ float abm_hi2<float>(float a, float b) { return fmaxf(a, b); }
template <> __device__ __forceinline__ double abm_lo2<double>(double a, double b) { return fmin(a, b); }
template <> __device__ __forceinline__ double abm_hi2<double>(double a, double b) { return fmax(a, b); }
template <typename T> __device__ __forceinline__ T abm_min(T a, T b) {
  if (abm_isnan(a)) return a;
  if (abm_isnan(b)) return b;
  return abm_lo2<T>(a, b);
}
template <typename T> __device__ __forceinline__ T abm_max(T a, T b) {
  if (abm_isnan(a)) return a;
  if (abm_isnan(b)) return b;
  return abm_hi2<T>(a, b);
}
// Welford accumulators as torch's var/std reduction keeps them in float32
// (mean, sum of squared deviations, count); abm_wf_first is one value added
// to the empty accumulator, abm_wf_merge the pairwise merge of Chan, Golub
// and LeVeque (1979) with its two multiply-adds fused, as torch's compiled
// merge does, and abm_wf_var the variance with correction c
struct abm_welford { float m, m2, nf; };
__device__ __forceinline__ abm_welford abm_wf_first(float x) {
  const float d = x - 0.0f;
  const float m = 0.0f + d / 1.0f;
  return {m, 0.0f + d * (x - m), 1.0f};
}
__device__ __forceinline__ abm_welford abm_wf_merge(abm_welford a, abm_welford b) {
  if (a.nf == 0.0f) return b;
  if (b.nf == 0.0f) return a;
  const float d = b.m - a.m;
  const float cnt = a.nf + b.nf;
  const float w = b.nf / cnt;
  const float t = d * d * a.nf;
  return {__builtin_fmaf(d, w, a.m), __builtin_fmaf(t, w, a.m2 + b.m2), cnt};
}
__device__ __forceinline__ float abm_wf_var(abm_welford a, float c) {
  const float div = a.nf > c ? a.nf - c : 0.0f;
  return a.m2 / div;
}
// the order torch's radix sort gives float32 keys: -0 as +0, then the bits
// mapped so that unsigned order is the float order (a NaN with its sign bit
// set below -inf, one without it above +inf)
__device__ __forceinline__ uint32_t abm_sort_key(float x) {
  uint32_t b = __float_as_uint(x);
  if (b == 0x80000000u) b = 0u;
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
template <typename T> __device__ __forceinline__ T abm_clamp_lo(T x, T lo) {
  if (abm_isnan(x)) return x;
  return abm_hi2<T>(x, lo);
}
template <typename T> __device__ __forceinline__ T abm_clamp_hi(T x, T hi) {
  if (abm_isnan(x)) return x;
  return abm_lo2<T>(x, hi);
}
"""

_RTC_TYPES = "".join(f"typedef __hip_internal::{t} {t};\n" for t in (
    "int8_t", "int16_t", "int32_t", "int64_t", "uint8_t", "uint16_t", "uint32_t", "uint64_t"))

_prelude_text: Optional[str] = None


def _prelude() -> str:
    global _prelude_text
    if _prelude_text is None:
        parts = []
        for path in _PRELUDE_PATHS:
            with open(path) as f:
                parts.append(f.read())
        _prelude_text = _RTC_TYPES + "\n".join(parts) + _HELPERS
    return _prelude_text


def _source(tr: _Trace, body_outputs: List[str], name: str) -> str:
    nw = max(len(tr.words), 1)
    body = "\n  ".join(tr.lines + body_outputs)
    return (f"{_prelude()}\nstruct AbmWords {{ unsigned long long w[{nw}]; }};\n"
            f"extern \"C\" __global__ void __launch_bounds__(256) {name}(const AbmWords a) {{\n"
            f"  const int64_t n = (int64_t)a.w[0];\n"
            f"  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;\n"
            f"  if (i >= n) return;\n  {body}\n}}\n")


# ---------------------------------------------------------------- compile cache
class _Kernel:
    def __init__(self, src: str, name: str):
        from ._lib import check, lib

        code, size = C.c_void_p(), C.c_size_t()
        check(lib().abm_rtc_compile(src.encode(), _ARCH.encode(), C.byref(code), C.byref(size)),
              "rtc_compile")
        try:
            h = C.c_void_p()
            check(lib().abm_rtc_load(code, name.encode(), C.byref(h)), "rtc_load")
        finally:
            lib().abm_rtc_code_free(code)
        self.handle = h
        self.name = name

    def launch(self, n: int, words: np.ndarray, stream) -> None:
        from ._lib import check, lib

        grid = max(1, (n + 255) // 256)
        check(lib().abm_rtc_launch(self.handle, grid, 256, words.ctypes.data, words.nbytes,
                                   stream), "rtc_launch")


_active: List["_Trace"] = []   # traces in progress (jnp constants become literals)


def active_trace() -> Optional["_Trace"]:
    """The trace in progress on this thread, if any."""
    return _active[-1] if _active else None


_kernels: Dict[str, _Kernel] = {}
_refused: Dict[Any, str] = {}
_launches = 0


def compile_source(src: str) -> None:
    """Compile ``src`` with hiprtc without loading it (a check that needs no
    GPU); raises on a compile error."""
    from ._lib import check, lib

    code, size = C.c_void_p(), C.c_size_t()
    check(lib().abm_rtc_compile(src.encode(), _ARCH.encode(), C.byref(code), C.byref(size)),
          "rtc_compile")
    lib().abm_rtc_code_free(code)


def _pack(tr: _Trace) -> np.ndarray:
    out = np.zeros(max(len(tr.words), 1), np.uint64)
    for j, (kind, v) in enumerate(tr.words):
        if kind == "ptr":
            out[j] = np.uint64(v.data_ptr())
        elif kind == "f64":
            out[j] = np.frombuffer(struct.pack("<d", v), np.uint64)[0]
        else:
            out[j] = np.frombuffer(struct.pack("<q", int(v)), np.uint64)[0]
    return out


def _wrap_model_state(tr: _Trace, model_state):
    if not isinstance(model_state, dict):
        return model_state

    def wrap(v):
        if isinstance(v, (bool, int, float)):
            return tr.python_param(v)
        on_device = isinstance(v, torch.Tensor) and (v.is_cuda or tr.allow_host)
        if on_device and v.dim() == 0 and v.dtype in _KIND:
            return tr.uniform_tensor(v)
        if (on_device and v.dim() == 1 and v.dtype in _KIND and 0 < v.shape[0] <= _MAX_VEC and
                v.is_contiguous()):
            return tr.uniform_vector(v)
        return v

    def agents_array(v):
        if isinstance(v, torch.Tensor) and (v.is_cuda or tr.allow_host) and v.dtype in _KIND \
                and v.dim() in (1, 2):
            return SymArray(tr, v)
        return v

    out = {}
    for k, v in model_state.items():
        if k == "env" and isinstance(v, dict):
            out[k] = {ek: wrap(ev) for ek, ev in v.items()}
        elif k == "time_step":
            out[k] = wrap(v)
        elif k.startswith("agents_") and isinstance(v, dict):
            out[k] = {vk: agents_array(vv) for vk, vv in v.items()}
        else:
            out[k] = v
    return out


def _signature(fn, states, model_state) -> Tuple:
    def sig(v):
        if isinstance(v, torch.Tensor):
            return ("t", v.dtype, tuple(v.shape))
        if isinstance(v, (bool, int, float)):
            return ("p", type(v).__name__)
        return ("o", type(v).__name__)

    ms = ()
    if isinstance(model_state, dict):
        ms = tuple(sorted((k, sig(v)) for k, v in model_state.items() if k != "env"))
        env = model_state.get("env")
        if isinstance(env, dict):
            ms += tuple(sorted((k, sig(v)) for k, v in env.items()))
    st = tuple(sorted((k, sig(v)) for k, v in (states or {}).items()))
    return (getattr(fn, "__qualname__", repr(fn)), id(getattr(fn, "__func__", fn)), st, ms)


def trace(n: int, key: torch.Tensor, mode: int, states: Optional[Dict[str, Any]], model_state,
          call, device=None):
    """Trace ``call(states_sym, model_state_sym, agent_key_sym)`` for n
    agents; returns (kernel name, source, trace, output tensors).  Raises
    Unsupported."""
    tr = _Trace(n, mode, allow_host=device is not None and torch.device(device).type == "cpu")
    tr.word("i64", n)
    kw = tr.word("ptr", key.contiguous() if key.dtype == torch.int32 else
                 key.to(torch.int32).contiguous())
    tr.lines.append(f"const uint32_t *kp = (const uint32_t *)a.w[{kw}];")
    tr.lines.append(f"const abm::Key ak = abm::split_at(abm::Key{{kp[0], kp[1]}}, n, i, "
                    f"{tr.mode});")
    sym_states = ({k: tr.agent_input(v) for k, v in states.items()}
                  if states is not None else None)
    _active.append(tr)
    try:
        out = call(sym_states, _wrap_model_state(tr, model_state), SymKey(tr, "ak"))
    finally:
        _active.pop()
    if not isinstance(out, dict):
        raise Unsupported("the function did not return a dict")
    stores, outs = [], {}
    for k, v in out.items():
        if isinstance(v, torch.Tensor) and v.dim() <= 1:
            v = tr.voperand(v)   # a constant array: the same value for every agent
        if isinstance(v, SymVec):
            t = torch.empty((n, len(v)), dtype=_TORCH[v.kind], device=device or key.device)
            w = tr.word("ptr", t)
            for j, e in enumerate(v.items):
                stores.append(f"(({_CT[v.kind]} *)a.w[{w}])[i * {len(v)}LL + {j}] = {e.name};")
            outs[k] = t
            continue
        if isinstance(v, Sym):
            s = v if not v.weak else v.as_array()
        elif isinstance(v, (bool, int, float)):
            s = tr.const(v).as_array()
        else:
            raise Unsupported(f"output {k!r} of type {type(v).__name__}")
        t = torch.empty((n,), dtype=_TORCH[s.kind], device=device or key.device)
        w = tr.word("ptr", t)
        stores.append(f"(({_CT[s.kind]} *)a.w[{w}])[i] = {s.name};")
        outs[k] = t
    digest = hashlib.sha1("\n".join(tr.lines + stores).encode()).hexdigest()[:16]
    name = f"abm_fused_{digest}"
    return name, _source(tr, stores, name), tr, outs


def _needs_grad(*trees) -> bool:
    def any_grad(v):
        if isinstance(v, torch.Tensor):
            return v.requires_grad
        if isinstance(v, dict):
            return any(any_grad(x) for x in v.values())
        if isinstance(v, (list, tuple)):
            return any(any_grad(x) for x in v)
        return False

    return any(any_grad(t) for t in trees)


def run_agents(fn, n: int, key: torch.Tensor, mode: int, states: Optional[Dict[str, Any]],
               model_state, call) -> Optional[Dict[str, torch.Tensor]]:
    """Run ``call(states_sym, model_state_sym, agent_key_sym)`` for all n
    agents as one generated kernel; None when the function is not covered
    (remembered for the function and input signature)."""
    if not _ENABLED or not isinstance(key, torch.Tensor) or not key.is_cuda:
        return None
    if torch.is_grad_enabled() and _needs_grad(states, model_state):
        return None      # the generated kernel is not differentiable: run vmapped
    sig = _signature(fn, states, model_state)
    if sig in _refused:
        return None
    try:
        name, src, tr, outs = trace(n, key, mode, states, model_state, call)
    except Unsupported as e:
        _refused[sig] = str(e)
        return None
    except Exception as e:  # noqa: BLE001 - anything the tracer cannot follow
        # runs vmapped, where the function executes for real (and raises
        # there if it is itself in error)
        _refused[sig] = f"{type(e).__name__}: {e}"
        return None
    kern = _kernels.get(src)
    if kern is None:
        kern = _Kernel(src, name)
        _kernels[src] = kern
    from ._lib import stream_handle

    global _launches
    kern.launch(n, _pack(tr), stream_handle())
    _launches += 1
    return outs


def stats() -> Dict[str, int]:
    """Fused launches so far and kernels compiled."""
    return {"launches": _launches, "kernels": len(_kernels), "refused": len(_refused)}


def refused() -> Dict[Any, str]:
    """Functions (and input signatures) that run on the vmapped path, with
    the reason."""
    return dict(_refused)
