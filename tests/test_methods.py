"""The elementwise DeconvolutionMethods on the GPU (mvr_compute_quotient / _final_values / _final_values_mul)
against the oracle, bit for bit: the update of psi keeps Java's float and double rounding (no FMA contraction),
so the GPU and the oracle agree exactly on every value, special values included (integral <= 0, NaN, inf,
weights 0 and 1, maxIntensity 0).  The same device functions run inside the fused x_update pass.

computeFinalValuesMul raises a double product to the power 1/V; the GPU's pow and the host's may differ in the
last bit of the double, so that case is bounded by one float ulp instead of bit equality.
"""
import numpy as np
import pytest

from mvr import hip

pytestmark = pytest.mark.gpu


def update_inputs(n, seed, special=True):
    rng = np.random.default_rng(seed)
    last = rng.uniform(1e-4, 800.0, n).astype(np.float32)
    last[::97] = np.float32(1e-4)
    integral = rng.uniform(0.2, 3.0, n).astype(np.float32)
    weight = rng.uniform(0.0, 1.0, n).astype(np.float32)
    weight[::7] = 0.0
    weight[::11] = 1.0
    if special:
        integral[::13] = 0.0
        integral[::17] = -1.5
        integral[::19] = np.nan
        integral[::23] = np.inf
        integral[::29] = np.float32(1e-30)
    return last, integral, weight


@pytest.mark.parametrize("lam,max_i", [(0.0, 300.0), (6e-4, 300.0), (0.006, 1234.5), (6e-4, 0.0)])
def test_final_values_bitwise(oracle, lam, max_i):
    n = (1 << 18) + 7                                # ragged: not a multiple of any work-group size
    last, integral, weight = update_inputs(n, 11)
    got, (gs, gm) = hip.compute_final_values(last, integral, weight, lam, 1e-4, max_i)
    want, (os_, om) = oracle.compute_final_values(last, integral, weight, lam, 1e-4, max_i)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.isnan(gm) == np.isnan(om) and np.isnan(gs) == np.isnan(os_)
    # finite data: maxChange exact, sumChange to double rounding of a different summation order
    last, integral, weight = update_inputs(n, 12, special=False)
    got, (gs, gm) = hip.compute_final_values(last, integral, weight, lam, 1e-4, max_i)
    want, (os_, om) = oracle.compute_final_values(last, integral, weight, lam, 1e-4, max_i)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    assert gm == om
    assert gs == pytest.approx(os_, rel=1e-12, abs=1e-8)


def test_final_values_empty_and_single():
    e = np.zeros(0, np.float32)
    out, st = hip.compute_final_values(e, e, e, 6e-4, 1e-4, 300.0)
    assert out.size == 0 and st == (0.0, -1.0)
    out, st = hip.compute_final_values([2.0], [1.5], [1.0], 0.0, 1e-4, 300.0)
    assert out[0] == np.float32(3.0) and st == (1.0, 1.0)


def test_quotient_bitwise(oracle):
    rng = np.random.default_rng(3)
    n = 300007
    img = rng.uniform(-2.0, 60.0, n).astype(np.float32)
    img[::9] = 0.0
    blurred = rng.uniform(0.01, 50.0, n).astype(np.float32)
    got = hip.compute_quotient(img, blurred)
    np.testing.assert_array_equal(got.view(np.uint32), oracle.compute_quotient(img, blurred).view(np.uint32))


@pytest.mark.parametrize("V", [1, 3, 8])
def test_final_values_mul(oracle, V):
    rng = np.random.default_rng(V)
    n = 200003
    last = rng.uniform(1e-4, 500.0, n).astype(np.float32)
    integ = rng.uniform(0.3, 3.0, (V, n)).astype(np.float32)
    integ[0, ::31] = 0.0
    w = rng.uniform(0.0, 0.6, (V, n)).astype(np.float32)
    w[:, ::5] = 0.0
    got, (gs, gm) = hip.compute_final_values_mul(last, integ, w, 6e-4, 1e-4, 250.0)
    want, (os_, om) = oracle.compute_final_values_mul(last, integ, w, 6e-4, 1e-4, 250.0)
    diff = np.abs(got.view(np.int32).astype(np.int64) - want.view(np.int32).astype(np.int64))
    print(f"V={V}: {np.count_nonzero(diff)} of {n} values differ, at most {diff.max()} ulp")
    if V == 1:
        assert diff.max() == 0                       # pow(x, 1) is exact
    assert diff.max() <= 1 and np.count_nonzero(diff) <= n // 1000
    assert gm == pytest.approx(om, rel=1e-6)
    assert gs == pytest.approx(os_, rel=1e-6, abs=1e-4)
