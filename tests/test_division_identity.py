"""The step kernels divide the curvature's finite differences by the launch constants 2 dt and
dt dt as float(double(x) * r) with r the FP64 reciprocal of the float denominator. That equals
the correctly rounded float quotient for every float x (the product's relative error, about
2^-52, is smaller than the distance of any float quotient from a rounding boundary, and no
quotient of two floats with a denominator that is not a power of two is a midpoint). Checked
here over random bit patterns and edge values for several time steps."""
import numpy as np
import pytest


@pytest.mark.parametrize("dt", [0.005, 0.01, 0.001, 60.0, 0.3, 1e-4, 7.0, 0.125, 1.0 / 3.0])
def test_reciprocal_product_is_correctly_rounded_division(dt):
    dt = np.float32(dt)
    rng = np.random.default_rng(int(dt * 1e6) % 1000)
    bits = rng.integers(0, 2**32, size=2_000_000, dtype=np.uint64).astype(np.uint32)
    x = bits.view(np.float32)
    edge = np.array([0.0, -0.0, np.inf, -np.inf, np.finfo(np.float32).max, np.finfo(np.float32).tiny,
                     np.float32(1e-45), -np.float32(1e-45), 1.0, -3.0], np.float32)
    x = np.concatenate([x[~np.isnan(x)], edge])
    for c in (np.float32(2) * dt, dt * dt):
        r = 1.0 / np.float64(c)
        with np.errstate(over="ignore", under="ignore", invalid="ignore"):
            want = x / c
            got = (x.astype(np.float64) * r).astype(np.float32)
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
