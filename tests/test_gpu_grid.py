"""GPU: the Grid's Moore-8 neighbourhood kernels (abm_grid_moore_sum_f32,
abm_grid_moore_counts_i8, Grid.moore_sum / moore_counts) against numpy
brute force over the eight shifted copies of the array (np.roll on a torus,
zero padding on a bounded grid), and against the Schelling program's own
first-step happiness."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

OFFSETS = [(-1, -1), (-1, 0), (-1, 1), (0, -1), (0, 1), (1, -1), (1, 0), (1, 1)]
SHAPES = [(1, 1), (1, 5), (2, 2), (3, 7), (33, 65), (512, 512), (1000, 37), (64, 2049)]


@pytest.fixture(scope="module")
def J():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import jaxabm_amd

    return jaxabm_amd


def _shifted(a, dr, dc, periodic):
    """b[r][c] = a[r+dr][c+dc] (wrapped, or 0 outside)."""
    if periodic:
        return np.roll(np.roll(a, -dr, 0), -dc, 1)
    out = np.zeros_like(a)
    R, C = a.shape
    rs, re = max(0, -dr), min(R, R - dr)
    cs, ce = max(0, -dc), min(C, C - dc)
    if rs < re and cs < ce:
        out[rs:re, cs:ce] = a[rs + dr:re + dr, cs + dc:ce + dc]
    return out


def _grid(J, periodic):
    class M(J.Model):
        pass

    return J.Grid(M(), (8, 8), periodic=periodic)


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("shape", SHAPES)
def test_moore_sum_against_shifted_copies(J, shape, periodic):
    import torch

    rs = np.random.default_rng(shape[0] * 7 + shape[1])
    x = rs.standard_normal(shape).astype(np.float32) * np.float32(1000.0)
    got = _grid(J, periodic).moore_sum(torch.from_numpy(x).cuda()).cpu().numpy()
    acc = np.zeros(shape, np.float64)
    for dr, dc in OFFSETS:   # the kernel's fixed order, float64
        acc = acc + _shifted(x, dr, dc, periodic).astype(np.float64)
    np.testing.assert_array_equal(got, acc.astype(np.float32))


@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("shape", SHAPES)
def test_moore_counts_against_shifted_copies(J, shape, periodic):
    import torch

    rs = np.random.default_rng(shape[0] * 11 + shape[1])
    cells = rs.choice(np.array([0, 1, 2, 3], np.int8), size=shape, p=[0.2, 0.4, 0.3, 0.1])
    same, occ = _grid(J, periodic).moore_counts(torch.from_numpy(cells).cuda())
    ws = np.zeros(shape, np.int32)
    wo = np.zeros(shape, np.int32)
    for dr, dc in OFFSETS:
        nb = _shifted(cells, dr, dc, periodic)
        wo += nb != 0
        ws += (nb != 0) & (nb == cells)
    np.testing.assert_array_equal(same.cpu().numpy(), ws)
    np.testing.assert_array_equal(occ.cpu().numpy(), wo)


def test_moore_counts_give_schelling_first_step_happiness(J):
    """The Schelling program's first recorded fraction happy (512 x 512
    torus, seed 0) from the seeded grid's Moore counts: happy iff
    8 same >= 3 occupied (isolated agents happy)."""
    import torch

    from jaxabm_amd.programs.schelling import SchellingEngine

    H = W = 512
    eng = SchellingEngine(H, W, seed=0)
    grid = torch.from_numpy(eng.grid().copy()).cuda()
    same, occ = _grid(J, True).moore_counts(grid)
    agents = grid != 0
    happy = ((8 * same >= 3 * occ) & agents).sum().item()
    hf, _ = eng.run(1)
    assert np.float32(happy) / np.float32(int(agents.sum().item())) == hf[0]
    cells, pos, typ, rng = O.schelling_init(H, W, 0, O.MODE_PARTITIONABLE)
    ohf, _ = O.schelling_run(H, W, cells, pos, typ, rng, 1, O.MODE_PARTITIONABLE)
    assert hf[0] == ohf[0]


def test_bad_shapes_are_refused(J):
    from jaxabm_amd._lib import AbmError, check, lib

    assert lib().abm_grid_moore_sum_f32(0, 5, 1, None, None, None) != 0
    with pytest.raises(AbmError):
        check(lib().abm_grid_moore_counts_i8(4, 4, 0, None, None, None, None), "counts")
