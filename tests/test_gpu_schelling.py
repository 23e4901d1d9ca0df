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
@pytest.mark.parametrize("H,W,steps,seed", [(64, 64, 30, 0), (100, 37, 20, 3), (5, 7, 10, 1),
                                            (128, 96, 12, 2), (1030, 1030, 3, 7)])
def test_schelling_matches_oracle(schelling, mode, H, W, steps, seed):
    """1030 x 1030: more than 1,024 blocks of 1,024 cells, the chunked
    empty-cell prefix."""
    eng = schelling.SchellingEngine(H, W, seed=seed, mode=mode)
    cells, pos, typ, rng = O.schelling_init(H, W, seed, mode)
    assert eng.n_agents == len(pos)
    np.testing.assert_array_equal(eng.grid().reshape(-1), cells)
    np.testing.assert_array_equal(eng.positions(), pos)
    hf, mv = eng.run(steps)
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
