"""GPU: Barabasi-Albert edge lists built on the device (abm_ba_edges_device,
network.ba_edges_device / ba_csr_device) against the oracle's sequential
builder (the degree-repeated node list grown one node at a time), entry for
entry, and the 1M-node graph of the SIR workload against its committed CSR
checksum."""
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


@pytest.mark.parametrize("n,m,seed", [(2, 1, 0), (3, 5, 1), (6, 5, 2), (7, 5, 3), (1000, 1, 4),
                                      (1000, 4, 5), (4096, 2, 6), (20_000, 8, 7),
                                      (150_000, 4, 8), (50_000, 16, 9)])
def test_device_edges_equal_oracle(N, n, m, seed):
    """Star-only graphs (n <= m + 1), m = 1 (trees), hub-heavy small graphs
    and larger m (more duplicate draws to skip)."""
    src, dst = N.ba_edges_device(n, m, seed)
    osrc, odst = O.ba_edges(n, m, O.prng_key(seed))
    np.testing.assert_array_equal(src.cpu().numpy(), osrc)
    np.testing.assert_array_equal(dst.cpu().numpy(), odst)


def test_device_graph_1M_matches_checksum_and_host_builder(N):
    """BA(1,000,000, m = 4, seed 0) -- the measured SIR graph: edges equal the
    host builder's, and the device CSR built from them has the committed
    checksum."""
    n, m = 1_000_000, 4
    src, dst = N.ba_edges_device(n, m, 0)
    hsrc, hdst = N.ba_edges(n, m, 0)
    np.testing.assert_array_equal(src.cpu().numpy(), hsrc)
    np.testing.assert_array_equal(dst.cpu().numpy(), hdst)
    g = N.ba_csr_device(n, m, 0)
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "sir_ba1M_m4_365steps.json")))
    assert g.nnz == gold["csr_nnz"]
    assert N.csr_checksum(g.row_ptr.cpu().numpy(), g.col[: g.nnz].cpu().numpy()) == \
        gold["csr_sha256"]


def test_bad_arguments(N):
    from jaxabm_amd._lib import AbmError

    with pytest.raises(AbmError, match="m must be"):
        N.ba_edges_device(10, 0, 0)
    src, dst = N.ba_edges_device(1, 3, 0)
    assert src.numel() == 0 and dst.numel() == 0
