"""GPU: the device build of abm_math.h (exp, log, log1p, pow in float32; exp,
log in float64) computes the same bits as the oracle's build of the same
source, over random bit patterns, dense ranges and special values.  The
oracle's functions are pinned against the mathematics in
tests/test_oracle_math.py."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

FN = {"exp": 0, "log": 1, "log1p": 2, "pow": 3}


@pytest.fixture(scope="module")
def dev():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd._lib import check, lib, ptr, stream_handle

    def f32(fn, x, y=None):
        xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).cuda()
        yt = torch.from_numpy(np.ascontiguousarray(y, np.float32)).cuda() if y is not None else None
        out = torch.empty_like(xt)
        check(lib().abm_math_f32(FN[fn], ptr(xt), ptr(yt), ptr(out), xt.numel(),
                                 stream_handle()), "abm_math_f32")
        return out.cpu().numpy()

    def f64(fn, x):
        xt = torch.from_numpy(np.ascontiguousarray(x, np.float64)).cuda()
        out = torch.empty_like(xt)
        check(lib().abm_math_f64(FN[fn], ptr(xt), ptr(out), xt.numel(), stream_handle()),
              "abm_math_f64")
        return out.cpu().numpy()

    return f32, f64


def _same(a, b):
    if a.dtype == np.float32:
        ua, ub = a.view(np.uint32), b.view(np.uint32)
    else:
        ua, ub = a.view(np.uint64), b.view(np.uint64)
    same = (ua == ub) | (np.isnan(a) & np.isnan(b))
    assert same.all(), (np.count_nonzero(~same), a[~same][:4], b[~same][:4])


def _inputs_f32():
    rng = np.random.default_rng(77)
    bits = rng.integers(0, 1 << 32, 4_000_000, dtype=np.uint64).astype(np.uint32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 1.0, -1.0, 1e-45, -1e-45,
                        2.0 ** -126, 88.72283, 88.72284, -103.97, -104.0, -0.5, 0.5],
                       np.float32)
    return np.concatenate([bits.view(np.float32), rng.uniform(-110, 90, 1_000_000).astype(np.float32),
                           rng.uniform(-1, 1, 1_000_000).astype(np.float32), special])


@pytest.mark.parametrize("fn", ["exp", "log", "log1p"])
def test_f32_functions_bit_identical(dev, fn):
    f32, _ = dev
    x = _inputs_f32()
    _same(f32(fn, x), O.math_f32(fn, x))


def test_f32_pow_bit_identical(dev):
    f32, _ = dev
    rng = np.random.default_rng(78)
    s = np.concatenate([rng.uniform(0, 1, 1_000_000), [0.0, 1.0, 1e-45, 0.5]]).astype(np.float32)
    y = np.concatenate([(1.0 / rng.uniform(1e-3, 1, 1_000_000)), [3.0, np.inf, 2.0, 100.0]]).astype(np.float32)
    _same(f32("pow", s, y), O.math_f32("pow", s, y))


@pytest.mark.parametrize("fn", ["exp", "log"])
def test_f64_functions_bit_identical(dev, fn):
    _, f64 = dev
    rng = np.random.default_rng(79)
    bits = rng.integers(0, np.iinfo(np.uint64).max, 2_000_000, dtype=np.uint64)
    x = np.concatenate([bits.view(np.float64), rng.uniform(-750, 720, 1_000_000),
                        [0.0, -0.0, np.inf, -np.inf, np.nan, 5e-324, 709.79, -745.2]])
    _same(f64(fn, x), O.math_f64(fn, x))


def test_argument_errors(dev):
    from jaxabm_amd._lib import lib

    assert lib().abm_math_f32(9, None, None, None, 4, None) != 0
    assert lib().abm_math_f32(3, None, None, None, 0, None) == 0
    assert lib().abm_math_f64(2, None, None, 4, None) != 0
