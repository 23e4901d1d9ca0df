"""GPU parity: wealth-exchange program against the oracle.

Integer wealth, the Gini/mean records and the float32 positions are
bit-exact (normal()'s erf_inv calls the shared abm_math.h log1p on both
sides)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]


@pytest.fixture(scope="module")
def wealth():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd.programs import wealth

    return wealth


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n,steps,seed", [(1000, 100, 0), (2, 10, 1), (3000, 20, 5), (1025, 7, 9),
                                          (1023, 12, 4)])
def test_wealth_matches_oracle(wealth, mode, n, steps, seed):
    eng = wealth.WealthEngine(n, seed=seed, mode=mode)
    x, y, w, rng = O.wealth_init(n, seed, mode)
    np.testing.assert_array_equal(eng.x.cpu().numpy(), x)
    np.testing.assert_array_equal(eng.w.cpu().numpy(), w)
    mean, gini = eng.run(steps)
    om, og = O.wealth_run(x, y, w, rng, steps, mode)
    np.testing.assert_array_equal(eng.w.cpu().numpy(), w)
    np.testing.assert_array_equal(mean, om)
    np.testing.assert_array_equal(gini, og)
    np.testing.assert_array_equal(eng.rng, rng)
    np.testing.assert_array_equal(eng.x.cpu().numpy(), x)
    np.testing.assert_array_equal(eng.y.cpu().numpy(), y)


def test_wealth_batch_equals_alone(wealth):
    """A batched launch gives each instance exactly its solo results."""
    sizes, seeds, steps = [1000, 5000, 37, 2048], [0, 1, 2, 3], [30, 12, 50, 25]
    solo = [wealth.WealthEngine(n, seed=s).run(k) for n, s, k in zip(sizes, seeds, steps)]
    engines = [wealth.WealthEngine(n, seed=s) for n, s in zip(sizes, seeds)]
    mixed = wealth.run_batch(engines, steps)
    for (m1, g1), (m2, g2) in zip(solo, mixed):
        np.testing.assert_array_equal(m1, m2)
        np.testing.assert_array_equal(g1, g2)


def _run_raw(wealth, eng, steps, max_n, draws=False):
    """One instance through abm_wealth_run_batch with an explicit max_n, with
    or without the precomputed-draw buffer."""
    import torch
    from jaxabm_amd._lib import check, lib, ptr, stream_handle

    keys = eng.key_table(steps)
    mean = torch.empty(steps, dtype=torch.float32, device=eng.device)
    gini = torch.empty(steps, dtype=torch.float32, device=eng.device)
    buf = eng.draw_buffer(steps) if draws else None
    desc = torch.from_numpy(eng.descriptor(keys, steps, mean, gini, buf).view(np.uint8)).to(eng.device)
    check(lib().abm_wealth_run_batch(ptr(desc), 1, max_n, eng.mode, stream_handle()), "run")
    return mean.cpu().numpy(), gini.cpu().numpy()


@pytest.mark.parametrize("draws", [False, True])
@pytest.mark.parametrize("n,max_n", [(1000, 0), (1000, 500), (5000, 8000), (8000, 8000),
                                     (300, 100)])
def test_wealth_lds_and_global_paths_agree(wealth, n, max_n, draws):
    """Every path gives identical records and state: draws computed in the
    run or grid-wide before it; small instances with state in registers;
    LDS-resident runs (n <= max_n <= 8000) and global-buffer runs (max_n = 0,
    or an instance above max_n)."""
    steps = 40
    a = wealth.WealthEngine(n, seed=3)
    b = wealth.WealthEngine(n, seed=3)
    ma, ga = a.run(steps)                     # precomputed draws, max_n = n
    mb, gb = _run_raw(wealth, b, steps, max_n, draws)
    np.testing.assert_array_equal(ma, mb)
    np.testing.assert_array_equal(ga, gb)
    for t in ("w", "x", "y"):
        np.testing.assert_array_equal(getattr(a, t).cpu().numpy(), getattr(b, t).cpu().numpy())


def test_wealth_global_path_matches_oracle(wealth):
    """An instance above the LDS capacity runs on its global buffers."""
    n, steps = 9000, 8
    eng = wealth.WealthEngine(n, seed=11)
    x, y, w, rng = O.wealth_init(n, 11, eng.mode)
    mean, gini = eng.run(steps)
    om, og = O.wealth_run(x, y, w, rng, steps, eng.mode)
    np.testing.assert_array_equal(eng.w.cpu().numpy(), w)
    np.testing.assert_array_equal(mean, om)
    np.testing.assert_array_equal(gini, og)


@pytest.mark.parametrize("n,steps,seed", [(8001, 30, 1), (50_000, 40, 2), (300_000, 12, 3)])
def test_grid_path_matches_oracle_and_one_workgroup_run(wealth, n, steps, seed):
    """Above 8,000 agents WealthEngine.run spreads each step over the GPU
    (abm_wealth_run_grid): the oracle's records and state, and the same bits
    as the one-workgroup run of the batch entry point on global buffers."""
    import os

    eng = wealth.WealthEngine(n, seed=seed)
    assert n > wealth.GRID_MIN_AGENTS
    mean, gini = eng.run(steps)
    x, y, w, rng = O.wealth_init(n, seed, O.MODE_PARTITIONABLE)
    om, og = O.wealth_run_threads(x, y, w, rng, steps, max(1, min(16, os.cpu_count() or 1)),
                                  O.MODE_PARTITIONABLE)
    np.testing.assert_array_equal(mean, om)
    np.testing.assert_array_equal(gini, og)
    np.testing.assert_array_equal(eng.w.cpu().numpy(), w)
    np.testing.assert_array_equal(eng.x.cpu().numpy(), x)
    np.testing.assert_array_equal(eng.y.cpu().numpy(), y)
    one = wealth.WealthEngine(n, seed=seed)
    m2, g2 = _run_raw(wealth, one, steps, 0, draws=True)
    np.testing.assert_array_equal(m2, mean)
    np.testing.assert_array_equal(g2, gini)


def test_grid_path_wealth_above_the_lds_bins(wealth):
    """A state where one agent holds 5,000 units (total wealth 15,000 <= n,
    the histogram's range): wealth values past the grid path's per-workgroup
    LDS bins go to the global histogram directly; the records and state equal
    the oracle's from the same state."""
    import os

    import torch

    n, steps, seed = 20_000, 25, 5
    eng = wealth.WealthEngine(n, seed=seed)
    x, y, w, rng = O.wealth_init(n, seed, O.MODE_PARTITIONABLE)
    w[:] = 0
    w[20:10020] = 1
    w[7] = 5000
    eng.w.copy_(torch.from_numpy(w.copy()).to(eng.w.device))
    mean, gini = eng.run(steps)
    om, og = O.wealth_run_threads(x, y, w, rng, steps, max(1, min(16, os.cpu_count() or 1)),
                                  O.MODE_PARTITIONABLE)
    assert w.max() > 1024
    np.testing.assert_array_equal(mean, om)
    np.testing.assert_array_equal(gini, og)
    np.testing.assert_array_equal(eng.w.cpu().numpy(), w)
