"""Bulk Network construction: Barabasi-Albert edge lists and undirected CSR.

The reference's ``Network`` grows an ``(E, 2)`` int32 edge list one
``jnp.concatenate`` per ``add_edge`` and answers ``get_neighbors`` with O(E)
masks (jaxabm/agentpy.py:545-615).  Here the edge list is built in bulk by the
native library and stored as CSR (``row_ptr[n+1]``, ``col[2E]``, rows
ascending), which is also what ``get_neighbors`` returns for undirected
networks (ascending unique ids).
"""
from __future__ import annotations

import numpy as np
import torch

from ._lib import check, lib, ptr, stream_handle


def ba_edges(n: int, m: int, graph_seed: int) -> tuple[np.ndarray, np.ndarray]:
    """Barabasi-Albert(n, m) edges from key ``PRNGKey(graph_seed)``."""
    E = int(lib().abm_ba_num_edges(int(n), int(m)))
    src = np.zeros(max(E, 1), np.int32)
    dst = np.zeros(max(E, 1), np.int32)
    key = np.array([0, int(graph_seed) & 0xFFFFFFFF], np.uint32)
    got = int(lib().abm_ba_edges(int(n), int(m), key, src, dst))
    if got != E:
        check(1, "ba_edges")
    return src[:E], dst[:E]


def ba_edges_device(n: int, m: int, graph_seed: int, device="cuda"
                    ) -> tuple[torch.Tensor, torch.Tensor]:
    """The Barabasi-Albert(n, m) edges of ``ba_edges`` built on the device
    (abm_ba_edges_device): the same (src, dst) int32 tensors, entry for
    entry."""
    E = int(lib().abm_ba_num_edges(int(n), int(m)))
    src = torch.empty(max(E, 1), dtype=torch.int32, device=device)
    dst = torch.empty(max(E, 1), dtype=torch.int32, device=device)
    open_ = torch.empty(1, dtype=torch.int32, device=device)
    key = np.array([0, int(graph_seed) & 0xFFFFFFFF], np.uint32)
    check(lib().abm_ba_edges_device(int(n), int(m), key, ptr(src), ptr(dst), ptr(open_),
                                    stream_handle()), "ba_edges_device")
    return src[:E], dst[:E]


def ba_csr_device(n: int, m: int, graph_seed: int, device="cuda") -> "DeviceCSR":
    """Barabasi-Albert(n, m) as an undirected CSR built entirely on the device
    (edges, then abm_csr_from_edges with both directions): the arrays of
    ``ba_csr``."""
    src, dst = ba_edges_device(n, m, graph_seed, device)
    return csr_from_edges(n, src, dst, symmetrize=True, device=device)


def csr_undirected(n: int, src: np.ndarray, dst: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    src = np.ascontiguousarray(src, np.int32)
    dst = np.ascontiguousarray(dst, np.int32)
    E = len(src)
    row_ptr = np.zeros(n + 1, np.int32)
    col = np.zeros(max(2 * E, 1), np.int32)
    check(lib().abm_csr_undirected(int(n), E, src if E else np.zeros(1, np.int32),
                                   dst if E else np.zeros(1, np.int32), row_ptr, col),
          "csr_undirected")
    return row_ptr, col[: 2 * E]


def node_chunks(row_ptr: np.ndarray, align: int, node_cost: int, target: int,
                max_nodes: int, heavy_deg: int = 1 << 30, heavy_cost: int = 0) -> np.ndarray:
    """Chunk pointer over contiguous agent ranges of about ``target`` cost
    (cost = node_cost + degree, plus heavy_cost for degree > heavy_deg)."""
    import ctypes as C

    rp = np.ascontiguousarray(row_ptr, np.int32)
    n = len(rp) - 1
    k = int(lib().abm_node_chunks(n, rp, int(align), int(node_cost), int(heavy_deg),
                                  int(heavy_cost), int(target), int(max_nodes), None))
    if k < 0:
        check(1, "node_chunks")
    out = np.zeros(k + 1, np.int32)
    lib().abm_node_chunks(n, rp, int(align), int(node_cost), int(heavy_deg), int(heavy_cost),
                          int(target), int(max_nodes), C.c_void_p(out.ctypes.data))
    return out


class DeviceCSR:
    """A graph resident in HBM as CSR (row_ptr int32[n+1], col int32[nnz]);
    host copies are made on first use."""

    def __init__(self, n: int, nnz: int, row_ptr: torch.Tensor, col: torch.Tensor,
                 host_row_ptr: np.ndarray | None = None, host_col: np.ndarray | None = None):
        self.n = int(n)
        self.nnz = int(nnz)
        self.row_ptr = row_ptr
        self.col = col
        self._host_row_ptr = host_row_ptr
        self._host_col = host_col

    @classmethod
    def from_host(cls, row_ptr: np.ndarray, col: np.ndarray, device="cuda") -> "DeviceCSR":
        rp = torch.from_numpy(np.ascontiguousarray(row_ptr, np.int32)).to(device)
        cl = torch.from_numpy(np.ascontiguousarray(col if len(col) else np.zeros(1, np.int32),
                                                   np.int32)).to(device)
        return cls(len(row_ptr) - 1, int(row_ptr[-1]), rp, cl, row_ptr, col)

    @property
    def host_row_ptr(self) -> np.ndarray:
        if self._host_row_ptr is None:
            self._host_row_ptr = self.row_ptr.cpu().numpy()
        return self._host_row_ptr

    @property
    def host_col(self) -> np.ndarray:
        if self._host_col is None:
            self._host_col = self.col[: self.nnz].cpu().numpy()
        return self._host_col

    def neighbors(self, v: int) -> np.ndarray:
        return self.host_col[self.host_row_ptr[v]:self.host_row_ptr[v + 1]].copy()

    def row(self, v: int) -> torch.Tensor:
        """Row v's entries on the device (an int32 copy)."""
        if not 0 <= v < self.n:
            return torch.zeros(0, dtype=torch.int32, device=self.col.device)
        lo, hi = (int(x) for x in self.row_ptr[v:v + 2].cpu())
        return self.col[lo:hi].clone()


CSR_SYMMETRIZE, CSR_DEDUP, CSR_INSERTION_ORDER = 1, 2, 4   # abm_csr_from_edges flags


def csr_from_edges(n: int, src, dst, symmetrize: bool = False, dedup: bool = False,
                   insertion_order: bool = False, device="cuda") -> DeviceCSR:
    """CSR built on the device from the edge-list entries (src[i], dst[i]):
    rows ascending; neighbours ascending, or in entry order with
    ``insertion_order``; ``symmetrize`` adds (dst, src) after every entry;
    ``dedup`` drops repeated entries.  src/dst: host arrays or device tensors."""
    import ctypes as C

    s = torch.as_tensor(np.asarray(src, np.int32) if not torch.is_tensor(src) else src)
    d = torch.as_tensor(np.asarray(dst, np.int32) if not torch.is_tensor(dst) else dst)
    s = s.to(device=device, dtype=torch.int32).contiguous().reshape(-1)
    d = d.to(device=device, dtype=torch.int32).contiguous().reshape(-1)
    if s.numel() != d.numel():
        raise ValueError("src and dst need the same length")
    m = s.numel()
    flags = ((CSR_SYMMETRIZE if symmetrize else 0) | (CSR_DEDUP if dedup else 0)
             | (CSR_INSERTION_ORDER if insertion_order else 0))
    cap = 2 * m if symmetrize else m
    row_ptr = torch.empty(n + 1, dtype=torch.int32, device=device)
    col = torch.empty(max(cap, 1), dtype=torch.int32, device=device)
    ws = torch.empty(max(int(lib().abm_csr_from_edges_workspace(m, flags)), 1), dtype=torch.uint8,
                     device=device)
    nnz = C.c_int64(0)
    check(lib().abm_csr_from_edges(int(n), m, ptr(s), ptr(d), flags, ptr(row_ptr), ptr(col),
                                   C.byref(nnz), ptr(ws), ws.numel(), stream_handle()),
          "csr_from_edges")
    return DeviceCSR(n, nnz.value, row_ptr, col)


def ba_csr(n: int, m: int, graph_seed: int) -> tuple[np.ndarray, np.ndarray]:
    src, dst = ba_edges(n, m, graph_seed)
    return csr_undirected(n, src, dst)


def csr_checksum(row_ptr: np.ndarray, col: np.ndarray) -> str:
    """SHA-256 of the CSR arrays as little-endian int32 (row_ptr, then col):
    names a graph without shipping its arrays."""
    import hashlib

    h = hashlib.sha256()
    h.update(np.ascontiguousarray(row_ptr, "<i4").tobytes())
    h.update(np.ascontiguousarray(col, "<i4").tobytes())
    return h.hexdigest()


def csr_gather_sum(g: DeviceCSR, x: torch.Tensor) -> torch.Tensor:
    """out[v] = sum of x[u] over the neighbours u of v (float32 result,
    float64 accumulation)."""
    x = x.to(device=g.row_ptr.device, dtype=torch.float32).contiguous()
    if x.numel() < g.n:
        raise ValueError("x needs one value per node")
    out = torch.empty(g.n, dtype=torch.float32, device=g.row_ptr.device)
    check(lib().abm_csr_gather_sum_f32(g.n, ptr(g.row_ptr), ptr(g.col), ptr(x), ptr(out),
                                       stream_handle()), "csr_gather_sum")
    return out


def csr_count(g: DeviceCSR, mask: torch.Tensor) -> torch.Tensor:
    """out[v] = number of neighbours u of v with mask[u] (int32, exact)."""
    m = mask.to(device=g.row_ptr.device, dtype=torch.uint8).contiguous()
    if m.numel() < g.n:
        raise ValueError("mask needs one value per node")
    out = torch.empty(g.n, dtype=torch.int32, device=g.row_ptr.device)
    check(lib().abm_csr_count_u8(g.n, ptr(g.row_ptr), ptr(g.col), ptr(m), ptr(out),
                                 stream_handle()), "csr_count")
    return out


def scatter_add(out: torch.Tensor, idx: torch.Tensor, val: torch.Tensor) -> torch.Tensor:
    """out[idx[i]] += val[i] in place (float32 or int32; idx checked on the
    host against out's length)."""
    idx = idx.to(device=out.device, dtype=torch.int32).contiguous()
    val = val.to(device=out.device, dtype=out.dtype).contiguous()
    if idx.numel() != val.numel():
        raise ValueError("idx and val differ in length")
    if idx.numel() and (int(idx.min()) < 0 or int(idx.max()) >= out.numel()):
        raise IndexError("scatter_add index out of range")
    fn = {torch.float32: lib().abm_scatter_add_f32,
          torch.int32: lib().abm_scatter_add_i32}.get(out.dtype)
    if fn is None:
        raise TypeError("scatter_add: float32 or int32 output")
    check(fn(idx.numel(), ptr(idx), ptr(val), ptr(out), stream_handle()), "scatter_add")
    return out
