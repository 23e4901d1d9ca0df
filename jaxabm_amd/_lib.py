"""ctypes binding of ``libjaxabm_hip.so`` (declared in ``include/jaxabm_amd.h``).

The engine has one execution path: the HIP library on a ROCm GPU.  Loading
fails loudly when the library is missing, and every compute entry point
refuses tensors that are not on the GPU; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

from .build import LIB_PATH

MODE_ORIGINAL = 0
MODE_PARTITIONABLE = 1

GAMMA, LOGGAMMA, BETA = 0, 1, 2   # abm_gamma_each kinds
MAX_COPY_ITEMS = 64                # abm_copy_items


class CopyItem(C.Structure):
    _fields_ = [("src", C.c_void_p), ("dst", C.c_void_p), ("bytes", C.c_int64),
                ("row_stride", C.c_int64), ("row_limit", C.c_int64)]

REDUCE_BATCH_MAX = 16              # abm_reduce_batch
WEALTH_GRID_ACC_WORDS = 1026       # abm_wealth_run_grid's acc scratch
(REDUCE_SUM_F32, REDUCE_MEAN_F32, REDUCE_SUM_I32, REDUCE_MEAN_I32,
 REDUCE_SUM_U8) = range(5)


class ReduceItem(C.Structure):
    _fields_ = [("x", C.c_void_p), ("n", C.c_int64), ("out", C.c_void_p), ("op", C.c_int32),
                ("reserved", C.c_int32)]

_lock = threading.Lock()
_lib = None

_p = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int32
_int = C.c_int
_f32 = C.c_float
_u32 = C.c_uint32
_sz = C.c_size_t
_u32arr = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_i32arr = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")


class AbmError(RuntimeError):
    """A non-zero status from the C-ABI, with the library's message."""


class SchellingGrid(C.Structure):
    _fields_ = [
        ("height", C.c_int32), ("width", C.c_int32), ("n_agents", C.c_int64),
        ("cells", C.c_void_p), ("id", C.c_void_p), ("claim", C.c_void_p),
        ("pos", C.c_void_p), ("type", C.c_void_p), ("target", C.c_void_p),
        ("empty", C.c_void_p), ("block_counts", C.c_void_p),
    ]


DEFFUANT_TAPE_LAST = 1  # abm_deffuant_problem.flags bit


class DeffuantProblem(C.Structure):
    _fields_ = [
        ("n", C.c_int64), ("batch", C.c_int32), ("steps", C.c_int32),
        ("ckpt_every", C.c_int32), ("pair_blocks", C.c_int32), ("stat_chunks", C.c_int32),
        ("flags", C.c_int32),
        ("perms", C.c_void_p), ("eps", C.c_void_p), ("mu", C.c_void_p), ("work", C.c_void_p),
        ("ckpt", C.c_void_p), ("traj", C.c_void_p), ("lam", C.c_void_p), ("acc", C.c_void_p),
        ("partial", C.c_void_p), ("sums", C.c_void_p), ("coef", C.c_void_p),
        ("target", C.c_void_p), ("stats", C.c_void_p), ("loss", C.c_void_p),
        ("grads", C.c_void_p), ("weights", C.c_void_p),
    ]


class ServeConfig(C.Structure):
    _fields_ = [
        ("hbm_budget", C.c_uint64), ("max_agents", C.c_int64), ("max_active", C.c_int32),
        ("prefetch", C.c_int32), ("max_inflight", C.c_int32), ("steps_per_tick", C.c_int32),
        ("prep_threads", C.c_int32), ("mode", C.c_int32), ("sir_chunk", C.c_int64),
        ("wealth_chunk", C.c_int64), ("sir_resident_max", C.c_int64),
        ("wealth_resident_max", C.c_int64), ("strict_fifo", C.c_int32), ("reserved", C.c_int32),
    ]


class ServeRequest(C.Structure):
    _fields_ = [
        ("kind", C.c_int32), ("steps", C.c_int32), ("n", C.c_int64), ("seed", C.c_int64),
        ("graph_seed", C.c_int64), ("m", C.c_int32), ("n_infected", C.c_int32),
        ("beta", C.c_float), ("gamma", C.c_float),
    ]


class TicketInfo(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("kind", C.c_int32), ("steps", C.c_int32), ("resident", C.c_int32),
        ("n", C.c_int64), ("nnz", C.c_int64), ("device_bytes", C.c_int64),
        ("result_bytes", C.c_int64), ("t_submit", C.c_double), ("t_ready", C.c_double),
        ("t_admit", C.c_double), ("t_start", C.c_double), ("t_done", C.c_double),
    ]


class ServeStats(C.Structure):
    _fields_ = [(f, C.c_int64) for f in ("ticks", "completed", "agent_steps", "peak_active",
                                         "peak_device_bytes", "outstanding")]


class SirGraph(C.Structure):
    _fields_ = [
        ("n", C.c_int64),
        ("nnz", C.c_int64),
        ("row_ptr", C.c_void_p),
        ("col", C.c_void_p),
        ("chunk_ptr", C.c_void_p),
        ("n_chunks", C.c_int64),
    ]


# Device-side descriptor layouts (numpy structured dtypes with C alignment).
WEALTH_INST = np.dtype([("n", np.int64), ("x", np.uint64), ("y", np.uint64), ("w", np.uint64),
                        ("inc", np.uint64), ("hist", np.uint64), ("keys", np.uint64),
                        ("steps", np.int32), ("mean_out", np.uint64), ("gini_out", np.uint64),
                        ("draws", np.uint64)],
                       align=True)


# name -> (restype, argtypes).  Every symbol declared in include/jaxabm_amd.h
# appears here; tests check the two lists agree.
SIGNATURES = {
    "abm_last_error": (C.c_char_p, []),
    "abm_version": (_int, []),
    "abm_threefry2x32": (_int, [_p, _p, _p, _p, _p, _i64, _p]),
    "abm_split": (_int, [_p, _i64, _p, _int, _p]),
    "abm_random_bits32": (_int, [_p, _i64, _p, _int, _p]),
    "abm_uniform_f32": (_int, [_p, _i64, _f32, _f32, _p, _int, _p]),
    "abm_normal_f32": (_int, [_p, _i64, _p, _int, _p]),
    "abm_bernoulli_u8": (_int, [_p, _i64, _f32, _p, _int, _p]),
    "abm_randint_i32": (_int, [_p, _i64, _i32, _i32, _p, _int, _p]),
    "abm_split_each": (_int, [_p, _i64, _i64, _p, _int, _p]),
    "abm_draw_each": (_int, [_p, _i64, _i64, _int, _f32, _f32, _p, _int, _p]),
    "abm_randint_each": (_int, [_p, _i64, _i64, _i32, _i32, _p, _int, _p]),
    "abm_gamma_each": (_int, [_p, _i64, _i64, _p, _p, _int, _p, _int, _p]),
    "abm_math_f32": (_int, [_int, _p, _p, _p, _i64, _p]),
    "abm_math_f64": (_int, [_int, _p, _p, _i64, _p]),
    "abm_permutation_workspace": (_sz, [_i64]),
    "abm_permutation_rounds": (_int, [_i64]),
    "abm_permutation": (_int, [_p, _i64, _p, _p, _sz, _int, _p]),
    "abm_sort_pairs_u32": (_int, [_p, _p, _i64, _p, _sz, _p]),
    "abm_host_split": (_int, [_u32arr, _i64, _u32arr, _int]),
    "abm_step_keys": (_int, [_u32arr, _i32, _i32, _i32, _i32, _int, _u32arr]),
    "abm_ba_num_edges": (_i64, [_i64, _i32]),
    "abm_ba_edges": (_i64, [_i64, _i32, _u32arr, _i32arr, _i32arr]),
    "abm_ba_edges_device": (_int, [_i64, _i32, _u32arr, _p, _p, _p, _p]),
    "abm_csr_undirected": (_int, [_i64, _i64, _i32arr, _i32arr, _i32arr, _i32arr]),
    "abm_node_chunks": (_i64, [_i64, _i32arr, _i32, _i64, _i32, _i64, _i64, _i32, _p]),
    "abm_csr_from_edges_workspace": (_sz, [_i64, _i32]),
    "abm_csr_from_edges": (_int, [_i64, _i64, _p, _p, _i32, _p, _p, C.POINTER(_i64), _p, _sz,
                                  _p]),
    "abm_sir_chunk_params": (_int, [C.POINTER(_i32), C.POINTER(_i64), C.POINTER(_i32),
                                    C.POINTER(_i64), C.POINTER(_i32), C.POINTER(_i32),
                                    C.POINTER(_i32)]),
    "abm_sir_bitmap_words": (_i64, [_i64]),
    "abm_sir_slots": (_int, []),
    "abm_sir_seed": (_int, [C.POINTER(SirGraph), _p, _i64, _p, _p, _p, _u32, _p]),
    "abm_sir_run": (_int, [C.POINTER(SirGraph), _p, _p, _p, _u32, _p, _i32, _f32, _f32,
                           _int, _p, _p, _p]),
    "abm_sir_step": (_int, [C.POINTER(SirGraph), _p, _p, _p, _p, _u32, _p, _f32, _f32,
                            _int, _p, _p, _p]),
    "abm_sir_reduce_counts": (_int, [_p, _i32, _p, _p]),
    "abm_sir_states": (_int, [_p, _i64, _p, _p]),
    "abm_wealth_init": (_int, [_p, _i64, _p, _p, _p, _int, _p]),
    "abm_wealth_run_batch": (_int, [_p, _i32, _i64, _int, _p]),
    "abm_wealth_run_grid": (_int, [_p, _p, _int, _p]),
    "abm_permutation_batch_workspace": (_sz, [_i64, _i32]),
    "abm_permutation_batch": (_int, [_p, _i64, _i32, _i64, _p, _p, _sz, _int, _p]),
    "abm_deffuant_init": (_int, [_p, _i64, _i32, _p, _int, _p]),
    "abm_deffuant_step": (_int, [_p, _p, _i64, _i32, _p, _p, _p, _i32, _p]),
    "abm_deffuant_forward": (_int, [C.POINTER(DeffuantProblem), _p]),
    "abm_deffuant_backward": (_int, [C.POINTER(DeffuantProblem), _p]),
    "abm_serve_default_config": (None, [C.POINTER(ServeConfig)]),
    "abm_serve_clock": (C.c_double, []),
    "abm_serve_create": (_int, [C.POINTER(ServeConfig), C.POINTER(C.c_void_p)]),
    "abm_serve_submit": (_int, [_p, C.POINTER(ServeRequest), C.POINTER(C.c_int64)]),
    "abm_serve_poll": (_int, [_p, _i64, C.POINTER(C.c_int32)]),
    "abm_serve_wait": (_int, [_p, _i64, C.c_double, C.POINTER(C.c_int32)]),
    "abm_serve_collect": (_int, [_p, _i64, _p, _sz, C.POINTER(TicketInfo)]),
    "abm_serve_drain": (_int, [_p, C.c_double]),
    "abm_serve_get_stats": (_int, [_p, C.POINTER(ServeStats)]),
    "abm_serve_destroy": (_int, [_p]),
    "abm_csr_gather_sum_f32": (_int, [_i64, _p, _p, _p, _p, _p]),
    "abm_csr_count_u8": (_int, [_i64, _p, _p, _p, _p, _p]),
    "abm_scatter_add_f32": (_int, [_i64, _p, _p, _p, _p]),
    "abm_scatter_add_i32": (_int, [_i64, _p, _p, _p, _p]),
    "abm_grid_moore_sum_f32": (_int, [_i64, _i64, _i32, _p, _p, _p]),
    "abm_grid_moore_counts_i8": (_int, [_i64, _i64, _i32, _p, _p, _p, _p]),
    "abm_schelling_block_count": (_i64, [_i64]),
    "abm_schelling_init": (_int, [C.POINTER(SchellingGrid), _p, _int, C.POINTER(_i32), _p]),
    "abm_schelling_run": (_int, [C.POINTER(SchellingGrid), _p, _i32, _int, _p, _p, _p, _p]),
    "abm_schelling_run_from": (_int, [C.POINTER(SchellingGrid), _p, _i32, _int, _p, _p, _p, _p,
                                       _p]),
    "abm_schelling_settled": (_int, [C.POINTER(SchellingGrid), _p, _i32, _p, _p, _p, _p]),
    "abm_unit_threshold": (_u32, [_f32]),
    "abm_fill_i32": (_int, [_p, _i64, _i32, _p]),
    "abm_graph_begin": (_int, [_p]),
    "abm_graph_end": (_int, [_p, C.POINTER(_p)]),
    "abm_graph_launch": (_int, [_p, _p]),
    "abm_graph_destroy": (_int, [_p]),
    "abm_graph_sanitize": (_int, [_p, _p]),
    "abm_copy_items": (_int, [_p, _i32, _p, _p]),
    "abm_step_counters": (_int, [_p, _p, _p]),
    "abm_commit_step": (_int, [_p, _i32, _p, _p, _p, _p, _i64, _i32, _p, _p]),
    "abm_reduce_counts_u8": (_int, [_p, _i64, _i32, _p, _p]),
    "abm_reduce_workspace": (_sz, [_i64]),
    "abm_mean_f32": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_mean_i32": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_gini_i32": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_sum_f32": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_sum_i32": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_sum_u8": (_int, [_p, _i64, _p, _p, _sz, _p]),
    "abm_reduce_batch": (_int, [_p, _i32, _p]),
    "abm_soft_hist_f32": (_int, [_p, _i64, _i32, C.c_double, _p, _p, _sz, _p]),
    "abm_gini_i32_noalloc": (_int, [_p, _i64, _p, _p]),
    "abm_rtc_compile": (_int, [C.c_char_p, C.c_char_p, C.POINTER(_p), C.POINTER(C.c_size_t)]),
    "abm_rtc_code_free": (_int, [_p]),
    "abm_rtc_load": (_int, [_p, C.c_char_p, C.POINTER(_p)]),
    "abm_rtc_unload": (_int, [_p]),
    "abm_rtc_launch": (_int, [_p, _u32, _u32, _p, C.c_size_t, _p]),
    "abm_memcpy_d2d": (_int, [_p, _p, _sz, _p]),
    "abm_event_timer": (_int, [_p, _int, C.POINTER(_f32)]),
}


def lib():
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(
                    f"{LIB_PATH} is missing: build it with `python -m jaxabm_amd.build` "
                    "(the engine has no CPU fallback)")
            L = C.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(L, name)
                fn.restype = res
                fn.argtypes = args
            _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().abm_last_error()
        msg = msg.decode() if msg else "unknown error"
        raise AbmError(f"{what or 'jaxabm_amd'} failed (status {rc}): {msg}")


def require_gpu() -> None:
    import torch

    if not torch.cuda.is_available():
        raise RuntimeError("jaxabm_amd needs a ROCm GPU: torch.cuda.is_available() is False")


def ptr(t) -> C.c_void_p:
    """Raw device pointer of a torch tensor (must be on the GPU)."""
    if t is None:
        return C.c_void_p(0)
    if not t.is_cuda:
        raise ValueError("jaxabm_amd kernels take GPU tensors only (got a CPU tensor)")
    return C.c_void_p(t.data_ptr())


def stream_handle(stream=None) -> C.c_void_p:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)
