"""CPU oracle: a restatement of the reference's agent-step semantics in C.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this package, as the checker or
as the timed CPU baseline.  The shipped engine (``jaxabm_amd``) never imports
it, and fails loudly instead of falling back to it.

What it restates (reference files are cited in ``oracle.c``):
  * JAX threefry2x32 PRNG construction in both stream modes
    (jaxabm/model.py:46,129,156,164,183; jaxabm/agent.py:115,156);
  * the JaxModel step schedule (jaxabm/model.py:118-216) applied to four
    model programs written against the reference API (wealth exchange,
    Schelling, network SIR, Deffuant);
  * record-time reductions, Gini after
    examples/models/advanced_economic_model.py:1798-1811.

Parity pins for every function live in ``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

MODE_ORIGINAL = 0
MODE_PARTITIONABLE = 1
HIST_BINS = 32
STATS_LEN = 2 + HIST_BINS

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile ``liboracle.so`` with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "all"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i64 = C.c_int64
_i32 = C.c_int32


def _declare(L):
    sig = {
        "orc_threefry2x32": (None, [C.c_uint32] * 4 + [_u32p]),
        "orc_prng_key": (None, [_i64, _u32p]),
        "orc_split": (None, [_u32p, _i64, _u32p, C.c_int]),
        "orc_random_bits32": (None, [_u32p, _i64, _u32p, C.c_int]),
        "orc_uniform_f32": (None, [_u32p, _i64, C.c_float, C.c_float, _f32p, C.c_int]),
        "orc_randint_i32": (None, [_u32p, _i64, _i32, _i32, _i32p, C.c_int]),
        "orc_math_f32": (None, [C.c_int, _f32p, C.c_void_p, _f32p, _i64]),
        "orc_math_f64": (None, [C.c_int, _f64p, _f64p, _i64]),
        "orc_erfinv_f32": (C.c_float, [C.c_float]),
        "orc_normal_f32": (None, [_u32p, _i64, _f32p, C.c_int]),
        "orc_gamma_f32": (None, [_u32p, _i64, _f32p, C.c_int, _f32p, C.c_int]),
        "orc_beta_f32": (None, [_u32p, _i64, _f32p, _f32p, _f32p, C.c_int]),
        "orc_permutation_rounds": (_i32, [_i64]),
        "orc_permutation": (None, [_u32p, _i64, _i32p, C.c_int]),
        "orc_ba_num_edges": (_i64, [_i64, _i32]),
        "orc_ba_edges": (_i64, [_i64, _i32, _u32p, _i32p, _i32p]),
        "orc_csr_undirected": (None, [_i64, _i64, _i32p, _i32p, _i32p, _i32p]),
        "orc_gini_i32": (C.c_float, [_i32p, _i64]),
        "orc_mean_i32": (C.c_float, [_i32p, _i64]),
        "orc_sir_init": (None, [_i64, _i32, _i64, C.c_int, _u8p, _u32p]),
        "orc_sir_run_threads": (None, [_i64, _i32p, _i32p, _u8p, _u32p, C.c_float, C.c_float,
                                       C.c_int32, C.c_int, _i32p, C.c_int]),
        "orc_sir_run": (None, [_i64, _i32p, _i32p, _u8p, _u32p, C.c_float,
                               C.c_float, _i32, C.c_int, _i32p]),
        "orc_trader_init": (None, [_i64, _i64, C.c_int, _f32p, _i32p, _u32p]),
        "orc_trader_run": (None, [_i64, _f32p, _i32p, _u32p, _i32, C.c_int]),
        "orc_wealth_init": (None, [_i64, _i64, C.c_int, _f32p, _f32p, _i32p, _u32p]),
        "orc_wealth_run": (None, [_i64, _f32p, _f32p, _i32p, _u32p, _i32,
                                  C.c_int, _f32p, _f32p]),
        "orc_wealth_run_threads": (None, [_i64, _f32p, _f32p, _i32p, _u32p, _i32,
                                          C.c_int, _f32p, _f32p, C.c_int]),
        "orc_schelling_init": (_i64, [_i32, _i32, _i64, C.c_int, _i8p, _i32p,
                                      _i8p, _u32p]),
        "orc_schelling_run": (None, [_i32, _i32, _i8p, _i32p, _i8p, _i64,
                                     _u32p, _i32, C.c_int, _f32p, _i32p]),
        "orc_schelling_run_threads": (None, [_i32, _i32, _i8p, _i32p, _i8p, _i64,
                                             _u32p, _i32, C.c_int, _f32p, _i32p, C.c_int]),
        "orc_deffuant_init": (None, [_i64, _i32, _i64, C.c_int, _f32p, _u32p]),
        "orc_deffuant_forward": (None, [_i64, _i32, _f32p, _f32p, _f32p, _u32p,
                                        _i32, C.c_int]),
        "orc_deffuant_stats": (None, [_i64, _i32, _f32p, _f64p]),
        "orc_deffuant_loss": (C.c_double, [_f64p, _f64p]),
        "orc_deffuant_grad": (None, [_i64, _i32, _f64p, _f64p, _f32p, _u32p,
                                     _i32, C.c_int, _f64p, _f64p, _f64p, _f64p]),
        "orc_deffuant_grad_threads": (None, [_i64, _i32, _f64p, _f64p, _f32p, _u32p,
                                     _i32, C.c_int, _f64p, _f64p, _f64p, _f64p, _i32]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _key(key) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(key, dtype=np.uint32).reshape(2))


# ---------------------------------------------------------------- shared math
MATH_F32 = {"exp": 0, "log": 1, "log1p": 2, "pow": 3}
MATH_F64 = {"exp": 0, "log": 1}


def math_f32(fn: str, x, y=None) -> np.ndarray:
    """Elementwise float32 exp / log / log1p / pow of abm_math.h."""
    x = np.ascontiguousarray(x, dtype=np.float32).ravel()
    out = np.empty_like(x)
    yp = None
    if fn == "pow":
        y = np.ascontiguousarray(np.broadcast_to(np.asarray(y, np.float32), x.shape))
        yp = y.ctypes.data
    lib().orc_math_f32(MATH_F32[fn], x, yp, out, x.size)
    return out


def math_f64(fn: str, x) -> np.ndarray:
    """Elementwise float64 exp / log of abm_math.h."""
    x = np.ascontiguousarray(x, dtype=np.float64).ravel()
    out = np.empty_like(x)
    lib().orc_math_f64(MATH_F64[fn], x, out, x.size)
    return out


# ---------------------------------------------------------------- PRNG
def threefry2x32(key, x0: int, x1: int):
    out = np.zeros(2, np.uint32)
    k = _key(key)
    lib().orc_threefry2x32(int(k[0]), int(k[1]), x0 & 0xFFFFFFFF, x1 & 0xFFFFFFFF, out)
    return int(out[0]), int(out[1])


def prng_key(seed: int) -> np.ndarray:
    out = np.zeros(2, np.uint32)
    lib().orc_prng_key(int(seed), out)
    return out


def split(key, num: int = 2, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(2 * num, np.uint32)
    lib().orc_split(_key(key), num, out, mode)
    return out.reshape(num, 2)


def random_bits(key, n: int, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.uint32)
    lib().orc_random_bits32(_key(key), n, out, mode)
    return out[:n]


def uniform(key, n: int, lo: float = 0.0, hi: float = 1.0,
            mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_uniform_f32(_key(key), n, lo, hi, out, mode)
    return out[:n]


def randint(key, n: int, lo: int, hi: int, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_randint_i32(_key(key), n, lo, hi, out, mode)
    return out[:n]


def erfinv(x: float) -> float:
    return float(lib().orc_erfinv_f32(x))


def normal(key, n: int, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_normal_f32(_key(key), n, out, mode)
    return out[:n]


def _params(v, n: int) -> np.ndarray:
    return np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (n,)))


def gamma(key, a, n: int, log_space: bool = False,
          mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    """gamma(key, a, (n,)) (loggamma with log_space), a scalar or (n,)."""
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_gamma_f32(_key(key), n, _params(a, max(n, 1)), int(log_space), out, mode)
    return out[:n]


def beta(key, a, b, n: int, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.float32)
    lib().orc_beta_f32(_key(key), n, _params(a, max(n, 1)), _params(b, max(n, 1)), out, mode)
    return out[:n]


def permutation_rounds(n: int) -> int:
    return int(lib().orc_permutation_rounds(n))


def permutation(key, n: int, mode: int = MODE_PARTITIONABLE) -> np.ndarray:
    out = np.zeros(max(n, 1), np.int32)
    lib().orc_permutation(_key(key), n, out, mode)
    return out[:n]


# ---------------------------------------------------------------- network
def ba_edges(n: int, m: int, key):
    E = int(lib().orc_ba_num_edges(n, m))
    src = np.zeros(max(E, 1), np.int32)
    dst = np.zeros(max(E, 1), np.int32)
    got = lib().orc_ba_edges(n, m, _key(key), src, dst)
    assert got == E, (got, E)
    return src[:E], dst[:E]


def csr_undirected(n: int, src: np.ndarray, dst: np.ndarray):
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(max(2 * len(src), 1), np.int32)
    lib().orc_csr_undirected(n, len(src), src, dst, row_ptr, col)
    return row_ptr, col[: 2 * len(src)]


def ba_csr(n: int, m: int, graph_seed: int):
    src, dst = ba_edges(n, m, prng_key(graph_seed))
    return csr_undirected(n, src, dst)


# ---------------------------------------------------------------- reductions
def gini_i32(v) -> np.float32:
    v = np.ascontiguousarray(v, np.int32)
    return np.float32(lib().orc_gini_i32(v, len(v)))


def mean_i32(v) -> np.float32:
    v = np.ascontiguousarray(v, np.int32)
    return np.float32(lib().orc_mean_i32(v, len(v)))


# ---------------------------------------------------------------- programs
def sir_init(n: int, n_seed_infected: int, seed: int, mode: int = MODE_PARTITIONABLE):
    state = np.zeros(n, np.uint8)
    rng = np.zeros(2, np.uint32)
    lib().orc_sir_init(n, n_seed_infected, seed, mode, state, rng)
    return state, rng


def sir_run(row_ptr, col, state, rng, beta: float, gamma: float, steps: int,
            mode: int = MODE_PARTITIONABLE):
    """Advance ``state``/``rng`` in place; returns counts[steps, 3]."""
    counts = np.zeros((max(steps, 1), 3), np.int32)
    lib().orc_sir_run(len(state), row_ptr, np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32),
                      state, rng, beta, gamma, steps, mode, counts.reshape(-1))
    return counts[:steps]


def sir_run_threads(row_ptr, col, state, rng, beta: float, gamma: float, steps: int,
                    threads: int, mode: int = MODE_PARTITIONABLE):
    """sir_run with each step's agents on ``threads`` OpenMP threads."""
    counts = np.zeros((max(steps, 1), 3), np.int32)
    lib().orc_sir_run_threads(len(state), row_ptr, np.ascontiguousarray(col, np.int32),
                              state, rng, beta, gamma, steps, mode, counts.reshape(-1),
                              int(threads))
    return counts[:steps]


def trader_init(n: int, seed: int, mode: int = MODE_PARTITIONABLE):
    """The generic trader model's initial x, w and model key (see oracle.c)."""
    x = np.zeros(n, np.float32)
    w = np.zeros(n, np.int32)
    rng = np.zeros(2, np.uint32)
    lib().orc_trader_init(n, seed, mode, x, w, rng)
    return x, w, rng


def trader_run(x, w, rng, steps: int, mode: int = MODE_PARTITIONABLE):
    lib().orc_trader_run(len(x), x, w, rng, steps, mode)
    return x, w


def wealth_init(n: int, seed: int, mode: int = MODE_PARTITIONABLE):
    x = np.zeros(n, np.float32)
    y = np.zeros(n, np.float32)
    w = np.zeros(n, np.int32)
    rng = np.zeros(2, np.uint32)
    lib().orc_wealth_init(n, seed, mode, x, y, w, rng)
    return x, y, w, rng


def wealth_run(x, y, w, rng, steps: int, mode: int = MODE_PARTITIONABLE):
    mean = np.zeros(max(steps, 1), np.float32)
    gini = np.zeros(max(steps, 1), np.float32)
    lib().orc_wealth_run(len(x), x, y, w, rng, steps, mode, mean, gini)
    return mean[:steps], gini[:steps]


def wealth_run_threads(x, y, w, rng, steps: int, threads: int, mode: int = MODE_PARTITIONABLE):
    """wealth_run with each step's per-agent draws on ``threads`` threads
    (identical results)."""
    mean = np.zeros(max(steps, 1), np.float32)
    gini = np.zeros(max(steps, 1), np.float32)
    lib().orc_wealth_run_threads(len(x), x, y, w, rng, steps, mode, mean, gini, int(threads))
    return mean[:steps], gini[:steps]


def schelling_init(height: int, width: int, seed: int, mode: int = MODE_PARTITIONABLE):
    nc = height * width
    cells = np.zeros(nc, np.int8)
    pos = np.zeros(nc, np.int32)
    typ = np.zeros(nc, np.int8)
    rng = np.zeros(2, np.uint32)
    na = int(lib().orc_schelling_init(height, width, seed, mode, cells, pos, typ, rng))
    return cells, pos[:na].copy(), typ[:na].copy(), rng


def schelling_run(height: int, width: int, cells, pos, typ, rng, steps: int,
                  mode: int = MODE_PARTITIONABLE):
    hf = np.zeros(max(steps, 1), np.float32)
    mv = np.zeros(max(steps, 1), np.int32)
    lib().orc_schelling_run(height, width, cells, pos, typ, len(pos), rng, steps,
                            mode, hf, mv)
    return hf[:steps], mv[:steps]


def schelling_run_threads(height: int, width: int, cells, pos, typ, rng, steps: int,
                          threads: int, mode: int = MODE_PARTITIONABLE):
    """schelling_run with each step's agent phase on ``threads`` OpenMP
    threads (same results)."""
    hf = np.zeros(max(steps, 1), np.float32)
    mv = np.zeros(max(steps, 1), np.int32)
    lib().orc_schelling_run_threads(height, width, cells, pos, typ, len(pos), rng, steps,
                                    mode, hf, mv, int(threads))
    return hf[:steps], mv[:steps]


def deffuant_init(n: int, batch: int, seed: int, mode: int = MODE_PARTITIONABLE):
    x = np.zeros(n * batch, np.float32)
    rng = np.zeros(2, np.uint32)
    lib().orc_deffuant_init(n, batch, seed, mode, x, rng)
    return x.reshape(n, batch), rng


def deffuant_forward(x, eps, mu, rng, steps: int, mode: int = MODE_PARTITIONABLE):
    n, batch = x.shape
    flat = np.ascontiguousarray(x, np.float32).reshape(-1)
    lib().orc_deffuant_forward(n, batch, np.ascontiguousarray(eps, np.float32),
                               np.ascontiguousarray(mu, np.float32), flat, rng,
                               steps, mode)
    return flat.reshape(n, batch)


def deffuant_stats(x) -> np.ndarray:
    n, batch = x.shape
    st = np.zeros(batch * STATS_LEN, np.float64)
    lib().orc_deffuant_stats(n, batch, np.ascontiguousarray(x, np.float32).reshape(-1), st)
    return st.reshape(batch, STATS_LEN)


def deffuant_loss(stats, target) -> float:
    return float(lib().orc_deffuant_loss(np.ascontiguousarray(stats, np.float64),
                                         np.ascontiguousarray(target, np.float64)))


def deffuant_grad_threads(x0, eps, mu, rng0, steps: int, target, threads: int,
                          mode: int = MODE_PARTITIONABLE):
    """deffuant_grad with the instances over ``threads`` OpenMP threads (same
    bits)."""
    n, batch = x0.shape
    loss = np.zeros(batch, np.float64)
    ge = np.zeros(batch, np.float64)
    gm = np.zeros(batch, np.float64)
    lib().orc_deffuant_grad_threads(n, batch, np.ascontiguousarray(eps, np.float64),
                                    np.ascontiguousarray(mu, np.float64),
                                    np.ascontiguousarray(x0, np.float32).reshape(-1),
                                    _key(rng0), steps, mode,
                                    np.ascontiguousarray(target, np.float64), loss, ge, gm,
                                    int(threads))
    return loss, ge, gm


def deffuant_grad(x0, eps, mu, rng0, steps: int, target, mode: int = MODE_PARTITIONABLE):
    """float64 loss and d loss / d(eps, mu) per instance (restated jax.grad)."""
    n, batch = x0.shape
    loss = np.zeros(batch, np.float64)
    ge = np.zeros(batch, np.float64)
    gm = np.zeros(batch, np.float64)
    lib().orc_deffuant_grad(n, batch, np.ascontiguousarray(eps, np.float64),
                            np.ascontiguousarray(mu, np.float64),
                            np.ascontiguousarray(x0, np.float32).reshape(-1),
                            _key(rng0), steps, mode,
                            np.ascontiguousarray(target, np.float64), loss, ge, gm)
    return loss, ge, gm
