"""The mode sum's Bessel evaluation (power series of J_{n+1}, J_{n+2} with a
per-waveform series length, three backward recurrence steps) against scipy's
J_n, through the library's host build of the same device function."""
import numpy as np
import pytest
from scipy import special


@pytest.fixture(scope="module")
def L():
    from emri_kludge_suite_amd import _lib
    return _lib


@pytest.mark.parametrize("e_bound", [0.05, 0.3, 0.5, 0.66])
def test_series_against_scipy(L, e_bound):
    J = np.zeros(5)
    worst_abs = 0.0
    for n in list(range(1, 70)) + [100, 128]:
        for e in np.linspace(0.0, e_bound, 23)[1:]:
            assert L.lib().aak_test_bessel_five(n, e, e_bound, L.dptr(J)) > 0
            k = np.arange(n - 2, n + 3)
            ref = special.jv(k, n * e)
            if n == 1:
                ref[0] = -special.jv(1, e)
            # the mode sum adds the five with O(1) weights: absolute error is what
            # matters; values that carry weight must also be right relatively
            worst_abs = max(worst_abs, np.max(np.abs(J - ref)))
            big = np.abs(ref) > 1e-6
            np.testing.assert_allclose(J[big], ref[big], rtol=1e-11, atol=0)
    assert worst_abs < 1e-15


def test_zero_eccentricity(L):
    J = np.zeros(5)
    L.lib().aak_test_bessel_five(1, 0.0, 0.3, L.dptr(J))
    np.testing.assert_array_equal(J, [0, 1, 0, 0, 0])
    L.lib().aak_test_bessel_five(2, 0.0, 0.3, L.dptr(J))
    np.testing.assert_array_equal(J, [1, 0, 0, 0, 0])
    L.lib().aak_test_bessel_five(7, 0.0, 0.3, L.dptr(J))
    np.testing.assert_array_equal(J, [0, 0, 0, 0, 0])


def test_out_of_range_is_refused(L):
    J = np.zeros(5)
    assert L.lib().aak_test_bessel_five(3, 0.7, 0.7, L.dptr(J)) < 0
    assert L.lib().aak_test_bessel_five(129, 0.1, 0.2, L.dptr(J)) < 0


@pytest.mark.parametrize("n_max", [6, 30, 50, 128])
def test_block_taylor_amplitudes_match_direct_evaluation(L, n_max):
    # the mode sum expands alpha_n, beta_n, gamma_n in e about each block's
    # mid eccentricity; within the block's range the expansion must agree with
    # the direct Bessel evaluation to rounding (the amplitudes scale with n)
    out = np.zeros(6)
    checked = 0
    for e_ref in (0.0, 1e-3, 0.1, 0.3, 0.5, 0.6):
        for dl in (0.0, 1e-6, 2e-5, 1e-4, 5e-4):
            if e_ref + dl > 0.66:
                continue
            for n in sorted({1, 2, 3, 5, max(1, n_max // 2), n_max}):
                for d in (-dl, -dl / 3, 0.0, dl / 2, dl):
                    if e_ref + d < 0:
                        continue
                    p = L.lib().aak_test_taylor_amplitudes(n_max, n, e_ref, dl, d, L.dptr(out))
                    if p < 0:
                        continue
                    scale = max(1.0, np.max(np.abs(out[3:])))
                    assert np.max(np.abs(out[:3] - out[3:])) <= 4e-15 * scale, (n_max, n, e_ref, dl, d, p)
                    checked += 1
    assert checked > 100
