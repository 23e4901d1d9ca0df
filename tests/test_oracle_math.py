"""Pins of the shared transcendental functions (jaxabm_amd/csrc/abm_math.h, as
compiled into the oracle) against the mathematics: float64 numpy as the
truth, ulp bounds, special values, exact identities and monotonicity.

The device kernels compile the same source; tests/test_gpu_math.py checks
that they compute the same bits.  A dropped or wrong polynomial term, a wrong
split of ln2, a wrong range-reduction step or a mishandled subnormal breaks
one of the bounds below."""
import math

import numpy as np
import pytest

import oracle as O

F32_INF = np.float32(np.inf)


def _ulps_f32(got, ref64):
    """|got - ref| in float32 ulps of the correctly rounded result."""
    r32 = ref64.astype(np.float32)
    sp = np.spacing(np.abs(r32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - ref64) / sp


def _rng():
    return np.random.default_rng(20261203)


def test_expf_within_one_ulp():
    rng = _rng()
    x = np.concatenate([
        rng.uniform(-104.0, 88.72, 1_000_000),
        rng.uniform(-1.0, 1.0, 500_000),
        rng.uniform(-1e-6, 1e-6, 50_000),
        rng.uniform(-103.97, -87.3, 200_000),   # subnormal results
    ]).astype(np.float32)
    got = O.math_f32("exp", x)
    ref = np.exp(x.astype(np.float64))
    err = _ulps_f32(got, ref)
    assert err.max() <= 1.0, err.max()
    assert (got == ref.astype(np.float32)).mean() > 0.85


def test_expf_special_values():
    big_ok = np.float32(float.fromhex("0x1.62e42ep6"))
    big_over = np.nextafter(big_ok, F32_INF)
    x = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, big_ok, big_over, -104.5, 1.0], np.float32)
    got = O.math_f32("exp", x)
    assert got[0] == 1.0 and got[1] == 1.0
    assert got[2] == np.inf and got[3] == 0.0 and np.isnan(got[4])
    assert np.isfinite(got[5]) and got[6] == np.inf
    assert got[7] == 0.0
    assert got[8] == np.float32(math.e)


def test_expf_monotone():
    x = np.linspace(-110, 88.7, 2_000_001, dtype=np.float32)
    got = O.math_f32("exp", x)
    assert np.all(np.diff(got) >= 0)


def test_logf_within_one_ulp():
    rng = _rng()
    bits = rng.integers(1, 0x7F800000, 1_000_000, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([bits.view(np.float32),
                        rng.uniform(0.5, 2.0, 500_000).astype(np.float32),
                        rng.uniform(0.999, 1.001, 100_000).astype(np.float32)])
    got = O.math_f32("log", x)
    ref = np.log(x.astype(np.float64))
    err = _ulps_f32(got, ref)
    assert err.max() <= 1.0, err.max()
    assert (got == ref.astype(np.float32)).mean() > 0.9


def test_logf_special_values_and_powers_of_two():
    x = np.array([1.0, 0.0, -0.0, -1.0, np.inf, np.nan, 1e-45, 2.0 ** -126], np.float32)
    got = O.math_f32("log", x)
    assert got[0] == 0.0
    assert got[1] == -np.inf and got[2] == -np.inf
    assert np.isnan(got[3]) and got[4] == np.inf and np.isnan(got[5])
    # subnormal and smallest normal inputs: k ln2 rounded
    np.testing.assert_array_equal(got[6:], np.float32([-149 * math.log(2), -126 * math.log(2)]))
    k = np.arange(-149, 128)
    pw = np.ldexp(np.float32(1.0), k).astype(np.float32)
    np.testing.assert_array_equal(O.math_f32("log", pw), (k * math.log(2)).astype(np.float32))


def test_logf_monotone_and_inverse_of_exp():
    x = np.linspace(1e-3, 50, 2_000_001, dtype=np.float32)
    got = O.math_f32("log", x)
    assert np.all(np.diff(got) >= 0)
    y = np.linspace(-20, 20, 100_001, dtype=np.float32)
    back = O.math_f32("log", O.math_f32("exp", y))
    np.testing.assert_allclose(back, y, rtol=0, atol=4e-6)


def test_log1pf_nearly_correctly_rounded():
    rng = _rng()
    x = np.concatenate([rng.uniform(-1, 1, 500_000), -rng.uniform(0, 1, 200_000),
                        10 ** rng.uniform(-30, 30, 200_000),
                        -(10 ** rng.uniform(-30, -0.01, 200_000))]).astype(np.float32)
    got = O.math_f32("log1p", x)
    ref = np.log1p(x.astype(np.float64))
    assert _ulps_f32(got, ref).max() <= 0.5 + 1e-6
    special = np.array([-1.0, 0.0, -0.0, 1e-45, -1e-45, np.inf, -2.0, np.nan], np.float32)
    g = O.math_f32("log1p", special)
    assert g[0] == -np.inf and g[1] == 0.0 and g[2] == 0.0 and np.signbit(g[2])
    assert g[3] == np.float32(1e-45) and g[4] == np.float32(-1e-45)
    assert g[5] == np.inf and np.isnan(g[6]) and np.isnan(g[7])


def test_powf_nearly_correctly_rounded():
    rng = _rng()
    # the gamma sampler's boost: (1 - uniform) ** (1 / alpha), alpha in (0, 1)
    s = rng.uniform(0, 1, 500_000).astype(np.float32)
    a = rng.uniform(0.01, 1.0, 500_000).astype(np.float32)
    y = (np.float32(1.0) / a).astype(np.float32)
    got = O.math_f32("pow", s, y)
    ref = np.power(s.astype(np.float64), y.astype(np.float64))
    normal = ref > 2.0 ** -126
    assert _ulps_f32(got[normal], ref[normal]).max() <= 0.5 + 1e-6
    assert np.all(np.abs(got[~normal].astype(np.float64) - ref[~normal]) <= 2.0 ** -149)
    g = O.math_f32("pow", np.float32([1.0, 0.0, 0.25, 4.0]), np.float32([7.5, 3.0, 0.5, 1.5]))
    np.testing.assert_array_equal(g, np.float32([1.0, 0.0, 0.5, 8.0]))


def test_float64_exp_log():
    rng = _rng()
    x = np.concatenate([rng.uniform(-745, 709.7, 500_000), rng.uniform(-1, 1, 200_000)])
    got = O.math_f64("exp", x)
    ref = np.exp(x)
    assert np.max(np.abs(got - ref) / np.spacing(ref)) <= 2.0
    y = 10 ** rng.uniform(-300, 300, 500_000)
    got = O.math_f64("log", y)
    ref = np.log(y)
    assert np.max(np.abs(got - ref) / np.spacing(np.abs(ref))) <= 2.0
    k = np.arange(-1074, 1024)
    got = O.math_f64("log", np.ldexp(1.0, k))
    want = k * math.log(2)   # the float64 product, itself within an ulp
    assert got[k == 0] == 0.0
    nz = k != 0
    assert np.max(np.abs(got[nz] - want[nz]) / np.spacing(np.abs(want[nz]))) <= 1.0
    sp = O.math_f64("exp", np.array([0.0, -np.inf, np.inf, 709.79, -745.2, np.nan]))
    assert sp[0] == 1.0 and sp[1] == 0.0 and sp[2] == np.inf and sp[3] == np.inf
    assert sp[4] == 0.0 and np.isnan(sp[5])


@pytest.mark.parametrize("x", [0.3, -0.7, 5.5, -60.0])
def test_sigmoid_gate_of_deffuant_uses_shared_exp(x):
    # the gate 1 / (1 + exp(-z)) at float32, as the program computes it
    z = np.float32(x)
    e = O.math_f32("exp", np.array([-z], np.float32))[0]
    g = np.float32(1.0) / (np.float32(1.0) + e)
    assert abs(float(g) - 1.0 / (1.0 + math.exp(-float(z)))) <= 2 * np.spacing(np.float32(g))
