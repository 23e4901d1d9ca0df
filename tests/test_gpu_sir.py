"""GPU parity: network SIR program against the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]


@pytest.fixture(scope="module")
def sir():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd.programs import sir

    return sir


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,m,steps,beta,gamma,seed", [
    (2000, 4, 60, 0.05, 0.1, 0),
    (5000, 2, 40, 0.3, 0.05, 3),
    (3000, 80, 25, 0.02, 0.2, 1),   # dense rows: every row is heavy
    (1000, 4, 30, 0.0, 0.1, 2),     # beta = 0: no new infection
    (1000, 4, 30, 1.0, 0.0, 2),     # beta = 1, gamma = 0
    (2000, 3, 300, 0.02, 0.01, 5),  # > 256 steps: risk marks wrap around
])
def test_sir_matches_oracle(sir, mode, n, m, steps, beta, gamma, seed):
    row_ptr, col = O.ba_csr(n, m, graph_seed=seed + 100)
    eng = sir.SIRNetworkEngine(n, csr=(row_ptr, col), beta=beta, gamma=gamma,
                               n_seed_infected=10, seed=seed, mode=mode)
    st, rng = O.sir_init(n, 10, seed, mode)
    np.testing.assert_array_equal(eng.state_host(), st)
    np.testing.assert_array_equal(eng.rng, rng)
    want = O.sir_run(row_ptr, col, st, rng, beta, gamma, steps, mode)
    got = eng.run(steps)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(eng.state_host(), st)
    np.testing.assert_array_equal(eng.rng, rng)
    assert (got.sum(axis=1) == n).all()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,m,steps,beta,gamma,seed", [
    # chunks of 8192 agents (the maximum tile): tiles with many rows to scan
    # take the CSR sweep (one and two rounds), tiles of more than 65535 CSR
    # entries take worklists with row_ptr from global memory, and worklist
    # overflow in those makes each thread scan its own rows
    (100_000, 4, 40, 0.3, 0.1, 11),
    (40_000, 2, 40, 0.5, 0.05, 12),
    (20_000, 4, 60, 0.1, 0.1, 14),
    (3000, 80, 20, 0.02, 0.2, 13),   # one tile of 467k entries, all rows heavy
])
def test_sir_tile_paths_match_oracle(sir, mode, n, m, steps, beta, gamma, seed):
    row_ptr, col = O.ba_csr(n, m, graph_seed=seed + 100)
    n0 = max(10, n // 50)
    eng = sir.SIRNetworkEngine(n, csr=(row_ptr, col), beta=beta, gamma=gamma,
                               n_seed_infected=n0, seed=seed, mode=mode, chunk_cost=1 << 40)
    assert eng.n_chunks == -(-n // 8192)
    st, rng = O.sir_init(n, n0, seed, mode)
    want = O.sir_run(row_ptr, col, st, rng, beta, gamma, steps, mode)
    got = eng.run(steps)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(eng.state_host(), st)
    assert want[:, 1].max() > n // 4   # the epidemic reaches most tiles


@pytest.mark.parametrize("mode", MODES)
def test_sir_two_runs_continue(sir, mode):
    """A second run() continues from the current state (jaxabm/model.py:227-235)."""
    n = 3000
    row_ptr, col = O.ba_csr(n, 4, graph_seed=9)
    eng = sir.SIRNetworkEngine(n, csr=(row_ptr, col), n_seed_infected=20, seed=4, mode=mode)
    st, rng = O.sir_init(n, 20, 4, mode)
    a = eng.run(13)
    b = eng.run(17)
    want = O.sir_run(row_ptr, col, st, rng, 0.05, 0.1, 30, mode)
    np.testing.assert_array_equal(np.concatenate([a, b]), want)


@pytest.mark.parametrize("mode", MODES)
def test_sir_full_size_matches_oracle(sir, mode):
    """The measured workload itself -- 1M-agent BA(m=4), 365 steps, 100 seed
    infections -- bit for bit against the oracle (its OpenMP path over the
    agents of each step, the same results as its serial path): every step's
    S/I/R counts and the final state; plus the invariants S + I + R = N, S
    never grows, R never shrinks."""
    import json
    import os

    from jaxabm_amd import network as N

    n, m = 1_000_000, 4
    # the graph the bench measures: the product's builder, checksummed
    row_ptr, col = N.ba_csr(n, m, graph_seed=0)
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "sir_ba1M_m4_365steps.json")))
    assert N.csr_checksum(row_ptr, col) == gold["csr_sha256"]
    orp, ocol = O.ba_csr(n, m, graph_seed=0)
    np.testing.assert_array_equal(row_ptr, orp)
    np.testing.assert_array_equal(col, ocol)
    assert row_ptr[-1] == 2 * m * (n - m)
    eng = sir.SIRNetworkEngine(n, csr=(row_ptr, col), seed=0, mode=mode)
    st, rng = O.sir_init(n, 100, 0, mode)
    threads = max(1, min(16, os.cpu_count() or 1))
    want = O.sir_run_threads(orp, ocol, st, rng, 0.05, 0.1, 365, threads, mode)
    got = np.concatenate([eng.run(12), eng.run(365 - 12)])   # a run continues the state
    np.testing.assert_array_equal(got, want)
    name = "partitionable" if mode == O.MODE_PARTITIONABLE else "original"
    np.testing.assert_array_equal(got, np.array(gold["counts"][name], np.int32))
    np.testing.assert_array_equal(eng.state_host(), st)
    assert (got.sum(axis=1) == n).all()
    assert (np.diff(got[:, 0]) <= 0).all()   # S never grows
    assert (np.diff(got[:, 2]) >= 0).all()   # R never shrinks
    assert got[-1, 1] == 0 and got[-1, 2] > n // 2


def test_seed_refuses_bad_chunk_tables(sir):
    """abm_sir_seed validates the chunk table the step kernel assumes (0 to
    n, ascending, 64-aligned inner boundaries, at most one tile per chunk)
    before launching anything."""
    import ctypes as C

    import torch

    from jaxabm_amd._lib import AbmError, SirGraph, check, lib, ptr, stream_handle

    n = 1000
    row_ptr, col = O.ba_csr(n, 4, graph_seed=2)
    eng = sir.SIRNetworkEngine(n, csr=(row_ptr, col), n_seed_infected=5, seed=0)
    perm = torch.arange(n, dtype=torch.int32, device="cuda")
    for bad in ([0, 500, 1000], [0, 512, 999], [0, 1000, 512], [64, 1000]):
        cp = torch.tensor(bad, dtype=torch.int32, device="cuda")
        g = SirGraph(n, eng.graph.nnz, eng.graph.row_ptr.data_ptr(), eng.graph.col.data_ptr(),
                     cp.data_ptr(), len(bad) - 1)
        rc = lib().abm_sir_seed(C.byref(g), ptr(perm), 5, ptr(eng.state), ptr(eng.mark),
                                ptr(eng.bitmaps), 0, stream_handle())
        with pytest.raises(AbmError, match="chunk_ptr"):
            check(rc, "abm_sir_seed")
