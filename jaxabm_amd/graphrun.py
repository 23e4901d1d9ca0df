"""Generic model steps replayed as one hipGraph per step.

The reference runs ``JaxModel.step`` as a Python loop of un-jitted
dispatches: three or more key splits, the vmapped agent updates, the env
callback and the metric reductions, each a separate launch per step
(jaxabm/model.py:146-216, run loop jaxabm/model.py:218-262).  For a model
written against this API with any per-agent code (torch ops through
``jaxabm_amd.numpy`` and ``jaxabm_amd.random``), ``Model.run(steps,
graph=True)`` turns that loop into graph replays:

1. the key schedule of the run (rng, sk = split(rng); per collection sk, ck =
   split(sk); with an env callback sk, uk = split(sk)) is computed on the
   host and uploaded once as a table;
2. step 1 runs eagerly through the same step body on the model's own env
   values (its result types say which env entries and metrics are Python
   scalars); step 2 runs eagerly on the device slots with torch's sync-debug
   mode set to error: a body that reads a device value on the host (item(),
   float(), nonzero, a copy from host memory) cannot be replayed, and the run
   continues eagerly;
3. the body of step 3 is captured into a torch CUDA graph (private memory
   pool).  ``abm_graph_sanitize`` then turns every memset node into a
   fill-kernel node and refuses graphs with host, event or allocation nodes.
   Memset nodes are replaced because in replays after the first they have
   given wrong values on this platform;
4. each replayed step reads its keys from row t of the table and the time
   step from a device counter.  It writes agent and env state back into the
   same buffers, and its metrics into row t of a device history.  So a graph
   of several step bodies in a row replays several steps: a run of at least
   BATCH_MIN_STEPS steps also captures STEPS_PER_LAUNCH bodies as one graph
   and replays it while that many steps are left (then the one-step graph),
   which saves most of a graph launch per step; shorter runs do not pay
   for the extra capture.

Results are those of the eager loop bit for bit (same kernels, same order).
Env values that are Python scalars become 0-d device tensors (int32,
float32, bool: JAX's default dtypes) for the run and Python scalars again
after it; ``model_state['time_step']`` is a 0-d int32 device tensor during a
replayed run.  Python values that a step computes from Python-side state are
constants of the graph; a step whose Python values differ between steps is
not replayed.  Capture evaluates the step's Python code once more than the
steps it runs (STEPS_PER_LAUNCH + 1 times more in a batched run).
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Dict, List, Optional

import numpy as np
import torch

from . import random as jr
from . import reductions
from ._lib import MAX_COPY_ITEMS, CopyItem, check, lib, stream_handle

_SCALARS = (bool, int, float)
# step bodies per replayed graph in runs of at least BATCH_MIN_STEPS steps
# (measured on a 10,000-agent model: about 24.6 us per replayed step one a
# launch, 18.9 us eight a launch; capturing the eight costs about 2 ms)
STEPS_PER_LAUNCH = 8
BATCH_MIN_STEPS = 400


class NotReplayable(Exception):
    """The step cannot be replayed as a graph; the run continues eagerly."""


def _scalar_tensor(v, device) -> Optional[torch.Tensor]:
    if isinstance(v, bool):
        return torch.full((), v, dtype=torch.bool, device=device)
    if isinstance(v, int) and -(1 << 31) <= v < (1 << 31):
        return torch.full((), v, dtype=torch.int32, device=device)
    if isinstance(v, float):
        return torch.full((), v, dtype=torch.float32, device=device)
    return None


def _same_python_value(a, b) -> bool:
    try:
        return type(a) is type(b) and bool(a == b)
    except Exception:  # noqa: BLE001 - values without a plain equality
        return a is b


class _Runner:
    def __init__(self, model, n: int):
        self.m = model
        self.n = n
        self.dev = jr._device()
        self.colls = list(model._agent_collections.items())
        self.has_env = model._update_state_fn is not None
        rng = np.array(model._rng, dtype=np.uint32).copy()
        tab = jr.step_key_table(rng, n, len(self.colls), self.has_env, False)
        self.rng_end = rng
        self.width = tab.shape[1] // 2
        self.table = torch.from_numpy(
            np.ascontiguousarray(tab).view(np.int32).reshape(n, self.width, 2).copy()).to(self.dev)
        self.row = torch.zeros((1,), dtype=torch.int64, device=self.dev)
        # the current step's keys: row 0 now, then written by each step's
        # commit for the next (no gather node in the replayed step)
        self.keys = self.table[0].clone()
        self.done = torch.zeros((1,), dtype=torch.int32, device=self.dev)
        self.ts = torch.full((), int(model._time_step), dtype=torch.int32, device=self.dev)
        self.static = {name: {k: v.clone() for k, v in c.states.items()} for name, c in self.colls}
        self.py_types: Dict[str, type] = {}
        self.env: Dict[str, Any] = dict(model._env_state)   # step 1 sees the model's own values
        self.env_python: Dict[str, Any] = {}
        self.hist: Dict[str, torch.Tensor] = {}
        self.metric_python: Dict[str, Any] = {}
        self.info: Dict[str, Any] = {"mode": "eager", "steps": n}

    # -------------------------------------------------------------- the body
    def body(self, time_step=None):
        keys = self.keys
        # the callbacks get copies of the env-slot dict: a callback that
        # writes into its env and returns it must not replace the fixed slot
        # tensors the captured graph reads and commit() writes
        current = {"time_step": self.ts if time_step is None else time_step,
                   "env": dict(self.env)}
        for name, _c in self.colls:
            current[f"agents_{name}"] = self.static[name]
        for i, (name, c) in enumerate(self.colls):
            c._states = self.static[name]
            c.update(current, keys[i], self.m.config)
        updated = {name: c.states for name, c in self.colls}
        env_new = dict(self.env)
        # the callbacks' whole-array sums and means go out as one launch
        with reductions.deferring():
            if self.has_env:
                env_new = self.m._update_state_fn(dict(self.env), updated, self.m._params,
                                                  keys[len(self.colls)])
                if not isinstance(env_new, dict):
                    raise NotReplayable("update_state_fn did not return a dict")
            metrics = (self.m._metrics_fn(env_new, updated, self.m._params)
                       if self.m._metrics_fn else {})
        return (updated, reductions.materialize(env_new),
                dict(reductions.materialize(metrics)))

    def _check(self, updated, env_new, metrics, first: bool):
        for name, _c in self.colls:
            out, st = updated[name], self.static[name]
            if set(out) != set(st):
                raise NotReplayable(f"collection {name!r} changed its variables")
            for k, t in out.items():
                if not isinstance(t, torch.Tensor) or t.shape != st[k].shape or t.dtype != st[k].dtype:
                    raise NotReplayable(f"variable {name}.{k} changed shape or dtype")
        if first:
            return
        if set(env_new) != set(self.env):
            raise NotReplayable("the env callback changed the env keys")
        for k, v in env_new.items():
            slot = self.env[k]
            if isinstance(v, torch.Tensor):
                if not isinstance(slot, torch.Tensor) or v.shape != slot.shape or v.dtype != slot.dtype:
                    raise NotReplayable(f"env[{k!r}] changed type, shape or dtype")
            elif k not in self.env_python or not _same_python_value(v, self.env_python[k]):
                raise NotReplayable(f"env[{k!r}] is a Python value that changes between steps")
        if set(metrics) != set(self.hist):
            raise NotReplayable("the metric names changed")
        for k, v in metrics.items():
            h = self.hist[k]
            if isinstance(v, torch.Tensor):
                if tuple(v.shape) != tuple(h.shape[1:]) or v.dtype != h.dtype:
                    raise NotReplayable(f"metric {k!r} changed shape or dtype")
            elif k not in self.metric_python or not _same_python_value(v, self.metric_python[k]):
                raise NotReplayable(f"metric {k!r} is a Python value that changes between steps")

    def commit(self, updated, env_new, metrics):
        """Write a step's results into the fixed buffers (recorded in the
        graph when called under capture).  A result that shares memory with
        one of the buffers (a variable returned unchanged under another name,
        an env entry moved to another key) is cloned first, so the order of
        the copies cannot matter."""
        pairs = []
        for name, c in self.colls:
            st = self.static[name]
            pairs += [(st[k], t) for k, t in updated[name].items()]
        env_fill = []
        for k, v in env_new.items():
            slot = self.env[k]
            if isinstance(v, torch.Tensor):
                pairs.append((slot, v))
            elif isinstance(slot, torch.Tensor):
                env_fill.append((slot, v))
        dst_storages = {d.untyped_storage().data_ptr() for d, _ in pairs}
        srcs = []
        for d, t in pairs:
            if t is d or (t.data_ptr() == d.data_ptr() and t.stride() == d.stride()):
                srcs.append(None)          # already in place
            elif t.untyped_storage().data_ptr() in dst_storages:
                srcs.append(t.clone())
            else:
                srcs.append(t)
        # contiguous same-dtype copies and history rows go to one launch
        items, keep = [], []

        def batched(src, dst, row_stride=0, row_limit=0) -> bool:
            if (len(items) >= MAX_COPY_ITEMS or not src.is_contiguous() or
                    not dst.is_contiguous() or src.dtype != dst.dtype or not src.is_cuda):
                return False
            nbytes = src.numel() * src.element_size()
            if nbytes != (row_stride or dst.numel() * dst.element_size()):
                return False
            items.append(CopyItem(src.data_ptr(), dst.data_ptr(), nbytes, row_stride, row_limit))
            keep.append(src)
            return True

        for (d, _), t in zip(pairs, srcs):
            if t is not None and not batched(t, d):
                d.copy_(t)
        for slot, v in env_fill:
            slot.fill_(v)
        for name, c in self.colls:
            c._states = self.static[name]
        for k, v in metrics.items():
            h = self.hist[k]
            if isinstance(v, torch.Tensor):
                row = v.reshape(h.shape[1:])
                if not batched(row, h, h[0].numel() * h.element_size(), h.shape[0]):
                    h.index_copy_(0, self.row, v.reshape(1, *h.shape[1:]))
            else:
                h.index_fill_(0, self.row, v)
        arr = (CopyItem * max(len(items), 1))(*items)
        check(lib().abm_commit_step(arr, len(items), self.row.data_ptr(), self.ts.data_ptr(),
                                    self.done.data_ptr(), self.table.data_ptr(), self.n,
                                    2 * self.width, self.keys.data_ptr(), stream_handle()),
              "commit_step")

    def first_step(self):
        """Step 1, eagerly, on the model's own env values and Python time
        step (exactly what ``Model.step`` computes): its result types say
        which env entries and metrics are Python scalars.  Defines the env
        slots (Python scalars become 0-d tensors) and the history buffers."""
        updated, env_new, metrics = self.body(time_step=int(self.m._time_step))
        self._check(updated, env_new, metrics, first=True)
        slots: Dict[str, Any] = {}
        for k, v in env_new.items():
            if isinstance(v, torch.Tensor):
                slots[k] = torch.empty_like(v)
            else:
                self.env_python[k] = v
                t = _scalar_tensor(v, self.dev) if isinstance(v, _SCALARS) else None
                if t is not None:
                    self.py_types[k] = type(v)
                    slots[k] = t
                else:
                    slots[k] = v
        self.env = slots
        for k, v in metrics.items():
            if isinstance(v, torch.Tensor):
                self.hist[k] = torch.empty((self.n, *v.shape), dtype=v.dtype, device=self.dev)
            elif isinstance(v, _SCALARS):
                t = _scalar_tensor(v, self.dev)
                if t is None:
                    raise NotReplayable(f"metric {k!r} is out of int32 range")
                self.metric_python[k] = v
                self.hist[k] = torch.empty((self.n,), dtype=t.dtype, device=self.dev)
            else:
                raise NotReplayable(f"metric {k!r} is neither a tensor nor a Python scalar")
        self.commit(updated, env_new, metrics)

    def second_step(self):
        """Step 2, eagerly, on the device slots, with synchronising calls
        raising: a step that reads a device value on the host cannot be
        replayed.  Nothing is written unless the step passes."""
        prev = torch.cuda.get_sync_debug_mode()
        torch.cuda.set_sync_debug_mode("error")
        try:
            updated, env_new, metrics = self.body()
        finally:
            torch.cuda.set_sync_debug_mode(prev)
        self._check(updated, env_new, metrics, first=False)
        self.commit(updated, env_new, metrics)

    def capture(self, steps: int = 1):
        """A graph of ``steps`` step bodies in a row, each committed before
        the next reads the fixed buffers."""
        from ._lib import check, lib

        g = torch.cuda.CUDAGraph(keep_graph=True)
        problem = None
        with torch.cuda.graph(g):
            for _ in range(steps):
                updated, env_new, metrics = self.body()
                try:
                    self._check(updated, env_new, metrics, first=False)
                except NotReplayable as e:
                    problem = e   # end the capture normally, then discard it
                    break
                self.commit(updated, env_new, metrics)
        if problem is not None:
            for name, c in self.colls:
                c._states = self.static[name]
            raise problem
        counts = (C.c_int32 * 4)()
        check(lib().abm_graph_sanitize(C.c_void_p(g.raw_cuda_graph()), counts), "graph_sanitize")
        g.instantiate()
        if steps == 1:
            self.info.update({"memsets_replaced": counts[0], "kernel_nodes": counts[1],
                              "memcpy_nodes": counts[2]})
        return g

    def finish(self, done: int):
        """Hand the state after ``done`` steps back to the model and append
        those steps' history rows."""
        m = self.m
        torch.cuda.current_stream().synchronize()
        for name, c in self.colls:
            c._states = self.static[name]
        env = {}
        for k, v in self.env.items():
            if k in self.py_types and isinstance(v, torch.Tensor):
                env[k] = self.py_types[k](v.item())
            else:
                env[k] = v
        m._env_state = env
        t0 = m._time_step
        m._time_step = t0 + done
        rng = np.array(m._rng, dtype=np.uint32).copy()
        if done == self.n:
            rng = self.rng_end
        else:
            jr.step_key_table(rng, done, len(self.colls), self.has_env, False)
        m._rng = rng
        if not m.config.track_history:
            return
        cols: Dict[str, Any] = {}
        for k, h in self.hist.items():
            if k in self.metric_python:
                kind = type(self.metric_python[k])
                cols[k] = [kind(v) for v in h[:done].tolist()]
            else:
                cols[k] = h
        for i in range(done):
            t = t0 + i + 1
            if t % m.config.collect_interval == 0:
                m._history.append({"time_step": t, "metrics": {k: v[i] for k, v in cols.items()}})


def _needs_grad(model) -> bool:
    from .fuse import _needs_grad as any_grad

    states = {name: c.states for name, c in model.agent_collections.items()}
    return any_grad(states, model._env_state, model._params)


def run_graph(model, n: int, between=None) -> Dict[str, Any]:
    """Run up to ``n`` steps of ``model``: steps 1 and 2 eagerly, steps
    3..n as replays of captured steps (STEPS_PER_LAUNCH a launch in runs of
    at least BATCH_MIN_STEPS steps while that many are left, else one).  Returns the run's info
    (``steps_done``: 0 when step 1 fails, 1 when step 2 cannot be replayed,
    2 when the capture is refused, n otherwise; ``mode``, ``reason``, node
    counts).  The model's state, env, keys and history end as after
    ``steps_done`` calls of ``model.step()``; the caller runs the remaining
    steps eagerly.  ``between`` (tests) is called after each replay but the
    last."""
    if torch.is_grad_enabled() and _needs_grad(model):
        # a replayed graph is cut from autograd: differentiable runs stay eager
        return {"mode": "eager", "steps": n, "steps_done": 0,
                "reason": "a parameter, env value or state requires grad"}
    r = _Runner(model, n)
    originals = {name: c.states for name, c in r.colls}
    try:
        r.first_step()
    except Exception as e:  # noqa: BLE001 - any failure: the plain loop runs instead
        for name, c in r.colls:
            c._states = originals[name]
        return {"mode": "eager", "steps": n, "steps_done": 0,
                "reason": f"step 1: {type(e).__name__}: {e}"}
    try:
        r.second_step()
    except Exception as e:  # noqa: BLE001
        r.finish(1)
        return {"mode": "eager", "steps": n, "steps_done": 1,
                "reason": f"step 2: {type(e).__name__}: {e}"}
    if n == 2:
        r.finish(2)
        return {"mode": "eager", "steps": n, "steps_done": 2, "reason": "two steps"}
    k = STEPS_PER_LAUNCH if n >= BATCH_MIN_STEPS and STEPS_PER_LAUNCH > 1 else 1
    try:
        g = r.capture()
        gk = r.capture(k) if k > 1 else None
    except Exception as e:  # noqa: BLE001
        r.finish(2)
        return {"mode": "eager", "steps": n, "steps_done": 2,
                "reason": f"capture: {type(e).__name__}: {e}"}
    left = n - 2
    while left > 0:
        if gk is not None and left >= k:
            gk.replay()
            left -= k
        else:
            g.replay()
            left -= 1
        if between is not None and left > 0:
            between()
    r.finish(n)
    del g, gk
    r.info.update({"mode": "graph", "steps_done": n, "steps_per_launch": k})
    return r.info
