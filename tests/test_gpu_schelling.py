"""GPU parity: Schelling program against the oracle, bit-exact (grid,
positions, fraction happy, moves)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]


@pytest.fixture(scope="module")
def schelling():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd.programs import schelling

    return schelling


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("H,W,steps,seed,chunk", [(64, 64, 30, 0, 8), (64, 64, 30, 0, 0),
                                                  (64, 64, 30, 0, 1), (100, 37, 20, 3, 3),
                                                  (5, 7, 10, 1, 2), (128, 96, 12, 2, 8),
                                                  (1030, 1030, 3, 7, 8)])
def test_schelling_matches_oracle(schelling, mode, H, W, steps, seed, chunk):
    """1030 x 1030: more than 1,024 blocks of 1,024 cells, the chunked
    empty-cell prefix.  chunk: steps per launch (0: one launch)."""
    eng = schelling.SchellingEngine(H, W, seed=seed, mode=mode)
    cells, pos, typ, rng = O.schelling_init(H, W, seed, mode)
    assert eng.n_agents == len(pos)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
    hf, mv = eng.run(steps, chunk)
    ohf, omv = O.schelling_run(H, W, cells, pos, typ, rng, steps, mode)
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(hf, ohf)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
    np.testing.assert_array_equal(eng.rng, rng)


def test_schelling_full_size(schelling):
    """The measured workload itself -- 512 x 512 torus, 200 steps -- exact
    against the oracle at every step (fraction happy, moves) and in the final
    grid and positions (the run split 10 + 190 steps: runs continue); the
    per-type counts are conserved and positions stay distinct.  The grid
    settles at step 24, so the last 176 steps are identity steps (the
    kernels exit after reading the previous step's happy count), and the
    second part starts with a computed step."""
    H = W = 512
    eng = schelling.SchellingEngine(H, W, seed=0)
    cells, pos, typ, rng = O.schelling_init(H, W, 0, O.MODE_PARTITIONABLE)
    counts = np.bincount(cells.astype(np.int64), minlength=3)
    ohf, omv = O.schelling_run(H, W, cells, pos, typ, rng, 200, O.MODE_PARTITIONABLE)
    hf, mv = eng.run(10)
    hf2, mv2 = eng.run(190)
    np.testing.assert_array_equal(np.concatenate([hf, hf2]), ohf)
    np.testing.assert_array_equal(np.concatenate([mv, mv2]), omv)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
    assert (np.bincount(cells.astype(np.int64), minlength=3) == counts).all()
    assert len(np.unique(pos)) == len(pos)
    assert ohf[-1] > ohf[0]
    settled = int(np.argmax(ohf == 1.0))
    assert 10 < settled < 100 and (ohf[settled:] == 1.0).all() and (omv[settled:] == 0).all()


@pytest.mark.parametrize("chunk", [8, 5, 1])
def test_schelling_settled_run_stops_launching(schelling, chunk):
    """512 x 512, 200 steps launched ``chunk`` steps at a time: once a launch
    ends with every agent happy, the steps after the launch queued behind it
    are recorded without a launch.  Records, grid and positions equal the
    oracle's and the one-launch run's; fewer than 200 steps were launched."""
    import torch

    H = W = 512
    S = 200
    cells, pos, typ, rng = O.schelling_init(H, W, 0, O.MODE_PARTITIONABLE)
    ohf, omv = O.schelling_run(H, W, cells, pos, typ, rng, S, O.MODE_PARTITIONABLE)
    settled = int(np.argmax(ohf == 1.0))
    eng = schelling.SchellingEngine(H, W, seed=0)
    keys = eng.key_table(S)
    counters = torch.full((S, 2), -7, dtype=torch.int32, device="cuda")
    happy = torch.full((S,), -7.0, dtype=torch.float32, device="cuda")
    moves = torch.full((S,), -7, dtype=torch.int32, device="cuda")
    launched = eng.run_settling(keys, S, counters, happy, moves, chunk)
    np.testing.assert_array_equal(happy.cpu().numpy(), ohf)
    np.testing.assert_array_equal(moves.cpu().numpy(), omv)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
    # launch i ends settled once its last step is at or past the first
    # all-happy step; launch i + 1 is then in flight
    first = -(-(settled + 1) // chunk) - 1
    assert launched == min(S, (first + 2) * chunk)
    assert launched < S


def test_schelling_settled_rejects_an_unsettled_count(schelling):
    """abm_schelling_settled writes the identity records only when the count
    it is given equals the number of agents; another count gives records
    marked invalid (-1, NaN, -1), not identity steps."""
    import ctypes as C

    import torch

    from jaxabm_amd._lib import check, lib, ptr, stream_handle

    eng = schelling.SchellingEngine(16, 16, seed=1)
    n = eng.n_agents
    for prev, ok in ((n, True), (n - 1, False)):
        pc = torch.tensor([prev, 0], dtype=torch.int32, device="cuda")
        counters = torch.zeros((5, 2), dtype=torch.int32, device="cuda")
        happy = torch.zeros(5, dtype=torch.float32, device="cuda")
        moves = torch.zeros(5, dtype=torch.int32, device="cuda")
        check(lib().abm_schelling_settled(C.byref(eng._g), ptr(pc), 5, ptr(counters), ptr(happy),
                                          ptr(moves), stream_handle()), "settled")
        c, h, m = counters.cpu().numpy(), happy.cpu().numpy(), moves.cpu().numpy()
        if ok:
            assert (c[:, 0] == n).all() and (c[:, 1] == 0).all()
            assert (h == 1.0).all() and (m == 0).all()
        else:
            assert (c == -1).all() and np.isnan(h).all() and (m == -1).all()


def test_schelling_near_the_grid_size_limit(schelling):
    """3,900 x 3,900 cells (14,854 blocks of 1,024, close to the 15,000-block
    limit): the agents kernel's largest prefix plus its static LDS fit the
    workgroup; 2 steps exact against the oracle."""
    H = W = 3900
    eng = schelling.SchellingEngine(H, W, seed=5)
    cells, pos, typ, rng = O.schelling_init(H, W, 5, O.MODE_PARTITIONABLE)
    assert eng.n_agents == len(pos)
    hf, mv = eng.run(2)
    ohf, omv = O.schelling_run(H, W, cells, pos, typ, rng, 2, O.MODE_PARTITIONABLE)
    np.testing.assert_array_equal(mv, omv)
    np.testing.assert_array_equal(hf, ohf)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
