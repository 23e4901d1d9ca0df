"""GPU parity: the HIP threefry kernels against the oracle, bit-exact."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

MODES = [O.MODE_ORIGINAL, O.MODE_PARTITIONABLE]
SIZES = [1, 2, 3, 10, 1001, 65536]


@pytest.fixture(scope="module")
def jr():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from jaxabm_amd import random as jr

    return jr


def _u32(t):
    return t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("mode", MODES)
def test_threefry_block_kat(jr, mode):
    import torch
    from jaxabm_amd._lib import lib, ptr, stream_handle, check

    kats = [((0, 0), (0, 0)), ((0xFFFFFFFF, 0xFFFFFFFF), (0xFFFFFFFF, 0xFFFFFFFF)),
            ((0x13198A2E, 0x03707344), (0x243F6A88, 0x85A308D3))]
    for k, x in kats:
        key = jr._as_key(np.array(k, np.uint32))
        x0 = torch.tensor(np.array([x[0]], np.uint32).view(np.int32), device="cuda")
        x1 = torch.tensor(np.array([x[1]], np.uint32).view(np.int32), device="cuda")
        y0 = torch.empty_like(x0)
        y1 = torch.empty_like(x1)
        check(lib().abm_threefry2x32(ptr(key), ptr(x0), ptr(x1), ptr(y0), ptr(y1), 1,
                                     stream_handle()))
        assert (int(_u32(y0)[0]), int(_u32(y1)[0])) == O.threefry2x32(k, *x)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("num", [1, 2, 3, 7, 1000])
def test_split(jr, mode, num):
    k = jr.PRNGKey(42)
    got = _u32(jr.split(k, num, mode=mode))
    np.testing.assert_array_equal(got, O.split(O.prng_key(42), num, mode))


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", SIZES)
def test_bits_uniform_normal(jr, mode, n):
    key = O.split(O.prng_key(7), 2, mode)[1]
    k = jr._as_key(key)
    np.testing.assert_array_equal(_u32(jr.bits(k, (n,), mode=mode)), O.random_bits(key, n, mode))
    np.testing.assert_array_equal(jr.uniform(k, (n,), mode=mode).cpu().numpy(),
                                  O.uniform(key, n, mode=mode))
    np.testing.assert_array_equal(jr.uniform(k, (n,), minval=-2.5, maxval=3.0, mode=mode).cpu().numpy(),
                                  O.uniform(key, n, -2.5, 3.0, mode=mode))
    got = jr.normal(k, (n,), mode=mode).cpu().numpy()
    want = O.normal(key, n, mode=mode)
    # erf_inv's log1p is the shared abm_math.h function and its sqrt is IEEE
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("lohi", [(0, 10), (0, 8), (-5, 1000), (3, 3), (0, 999), (-2**31, 2**31 - 1)])
def test_randint(jr, mode, lohi):
    key = O.prng_key(3)
    lo, hi = lohi
    got = jr.randint(jr._as_key(key), (2049,), lo, hi, mode=mode).cpu().numpy()
    np.testing.assert_array_equal(got, O.randint(key, 2049, lo, hi, mode))


@pytest.mark.parametrize("mode", MODES)
def test_bernoulli(jr, mode):
    key = O.prng_key(11)
    got = jr.bernoulli(jr._as_key(key), 0.3, (5001,), mode=mode).cpu().numpy()
    want = O.uniform(key, 5001, mode=mode) < np.float32(0.3)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("n", [1, 2, 10, 1000, 65536, 300001])
def test_permutation(jr, mode, n):
    key = O.prng_key(5)
    got = jr.permutation(jr._as_key(key), n, mode=mode).cpu().numpy()
    np.testing.assert_array_equal(got, O.permutation(key, n, mode))


@pytest.mark.parametrize("mode", MODES)
def test_each_variants(jr, mode):
    keys = O.split(O.prng_key(9), 37, mode)
    k = jr._as_key(keys)
    got = _u32(jr.split_each(k, 3, mode=mode))
    want = np.stack([O.split(kk, 3, mode) for kk in keys])
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(jr.uniform_each(k, (), mode=mode).cpu().numpy(),
                                  np.array([O.uniform(kk, 1, mode=mode)[0] for kk in keys]))
    np.testing.assert_array_equal(_u32(jr.bits_each(k, (5,), mode=mode)),
                                  np.stack([O.random_bits(kk, 5, mode) for kk in keys]))
    np.testing.assert_array_equal(jr.randint_each(k, (), 0, 17, mode=mode).cpu().numpy(),
                                  np.array([O.randint(kk, 1, 0, 17, mode)[0] for kk in keys]))
    np.testing.assert_array_equal(jr.normal_each(k, (2,), mode=mode).cpu().numpy(),
                                  np.stack([O.normal(kk, 2, mode) for kk in keys]))
