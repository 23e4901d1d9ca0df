"""GPU: the Network CSR built on the device (abm_csr_from_edges) against the
oracle's host builder and numpy, and the facade's get_neighbors against the
reference's formula (jaxabm/agentpy.py:584-615) evaluated on the same edge
list."""
import json
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def N():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd import network

    return network


def _host(g):
    return g.row_ptr.cpu().numpy(), g.col[: g.nnz].cpu().numpy()


def test_ba_1M_device_csr_equals_oracle_and_checksum(N):
    """BA(1,000,000, m = 4): every edge in both directions, rows and
    neighbours ascending: the same arrays as the oracle's host builder, and
    the committed checksum of the measured SIR graph."""
    n, m = 1_000_000, 4
    src, dst = O.ba_edges(n, m, O.prng_key(0))
    g = N.csr_from_edges(n, src, dst, symmetrize=True)
    rp, col = _host(g)
    orp, ocol = O.csr_undirected(n, src, dst)
    np.testing.assert_array_equal(rp, orp)
    np.testing.assert_array_equal(col, ocol)
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "sir_ba1M_m4_365steps.json")))
    assert g.nnz == gold["csr_nnz"]
    assert N.csr_checksum(rp, col) == gold["csr_sha256"]


def _want(n, s, d, sym, dedup, order):
    if sym:
        s, d = np.stack([s, d], 1).reshape(-1), np.stack([d, s], 1).reshape(-1)
    if order:
        idx = np.argsort(s, kind="stable")
        rs, cs = s[idx], d[idx]
    else:
        key = s.astype(np.int64) * (1 << 32) + d.astype(np.int64)
        key = np.unique(key) if dedup else np.sort(key)
        rs, cs = (key >> 32).astype(np.int32), (key & 0xFFFFFFFF).astype(np.int32)
    rp = np.zeros(n + 1, np.int64)
    np.add.at(rp, rs.astype(np.int64) + 1, 1)
    return np.cumsum(rp).astype(np.int32), cs.astype(np.int32)


@pytest.mark.parametrize("n,m,seed", [(1, 5, 0), (10, 60, 1), (5000, 40_000, 2),
                                      (200_000, 300_000, 3), (7, 0, 4), (3000, 9000, 5)])
@pytest.mark.parametrize("sym,dedup,order", [(False, False, False), (True, False, False),
                                             (False, True, False), (True, True, False),
                                             (False, False, True), (True, False, True)])
def test_edge_lists_against_numpy(N, n, m, seed, sym, dedup, order):
    """Duplicates, self-loops, isolated rows (every third node id unused in
    the last case), empty lists; rows of both small and large degree."""
    rs = np.random.default_rng(seed)
    s = rs.integers(0, n, m).astype(np.int32)
    d = rs.integers(0, n, m).astype(np.int32)
    s[: m // 5] = rs.integers(0, n // 1000 + 1, m // 5)   # a few hub rows
    if n == 3000:
        s, d = (s // 3) * 3, (d // 3) * 3
    g = N.csr_from_edges(n, s, d, symmetrize=sym, dedup=dedup, insertion_order=order)
    rp, col = _host(g)
    wrp, wcol = _want(n, s, d, sym, dedup, order)
    np.testing.assert_array_equal(rp, wrp)
    np.testing.assert_array_equal(col, wcol)
    assert g.nnz == len(wcol) == rp[-1]


def test_bad_arguments(N):
    import torch

    from jaxabm_amd._lib import AbmError

    with pytest.raises(AbmError, match="out of range"):
        N.csr_from_edges(10, np.array([0, 10], np.int32), np.array([1, 2], np.int32))
    with pytest.raises(AbmError, match="out of range"):
        N.csr_from_edges(10, np.array([0, 3], np.int32), np.array([-1, 2], np.int32))
    with pytest.raises(AbmError, match="DEDUP"):
        N.csr_from_edges(10, np.array([0], np.int32), np.array([1], np.int32), dedup=True,
                         insertion_order=True)
    g = N.csr_from_edges(4, torch.zeros(0, dtype=torch.int32), torch.zeros(0, dtype=torch.int32))
    assert g.nnz == 0 and g.row_ptr.cpu().tolist() == [0, 0, 0, 0, 0]


@pytest.mark.parametrize("directed", [False, True])
def test_facade_get_neighbors_match_reference_formula(N, directed):
    """Network.get_neighbors for every agent equals the reference's masks
    over the stored edge list: directed, the targets of the agent's entries
    in insertion order; undirected, the sorted unique ids of both directions
    (jaxabm/agentpy.py:584-615)."""
    import jaxabm_amd as J

    class M(J.Model):
        pass

    rs = np.random.default_rng(11 if directed else 12)
    net = J.Network(M(), directed=directed)
    pairs = rs.integers(0, 40, (300, 2))
    for a, b in pairs[:150]:
        net.add_edge(int(a), int(b))
    net.add_edges(pairs[150:, 0], pairs[150:, 1])
    e = net.edges().cpu().numpy()
    s, d = e[:, 0], e[:, 1]
    for a in range(42):
        got = net.get_neighbors(a).cpu().numpy()
        if directed:
            want = d[s == a]
        else:
            want = np.unique(np.concatenate([d[s == a], s[d == a]]))
        np.testing.assert_array_equal(got, want.astype(np.int32))
