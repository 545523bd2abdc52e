"""GPU parity of the inner-product stage (AAK_manager.cu:352-446,
kernel.cu:557-567): real FFT of both channels, PSD weighting and the two
sums <h|d> = 4 sum Re(conj(h w) . d) and <h|h> = 4 sum |h w|^2 over the
one-sided spectrum, DC and Nyquist bins included.

The reference is numpy's pocketfft on the same seeded inputs.  Power-of-two
lengths from 2^8 to 2^23 run the fused four-step FFT (every radix-4/radix-2
stage mix and both column-block widths); other lengths run rocFFT.  The
tolerance, 1e-12 relative, is three orders below the 1e-9 the likelihood
must meet.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12


def _case(N, seed, monkeypatch=None, force_rocfft=False):
    from emri_kludge_suite_amd import GPUAAK
    if monkeypatch is not None:
        if force_rocfft:
            monkeypatch.setenv("AAK_FFT", "rocfft")
        else:
            monkeypatch.delenv("AAK_FFT", raising=False)
    print("case fft N=%d seed=%d force_rocfft=%d" % (N, seed, force_rocfft), flush=True)
    rng = np.random.default_rng(seed)
    nb = N // 2 + 1
    h = rng.standard_normal((2, N))
    w = rng.uniform(0.5, 2.0, (2, nb))
    d = (rng.standard_normal((2, nb)) + 1j * rng.standard_normal((2, nb))) * w
    g = GPUAAK(1.0, max(N // 10, 16), N, 100.0, 10.0, True, False)
    g.input_data(d[0], d[1], w[0], w[1])
    (dh, hh), spec, fused = g.inner_product_stage(h[0], h[1])
    H = np.fft.rfft(h, axis=1) * w
    dh_ref = 4.0 * np.sum(np.real(np.conj(H) * d))
    hh_ref = 4.0 * np.sum(np.abs(H) ** 2)
    err_spec = float(np.sqrt(np.sum(np.abs(spec - H) ** 2) / np.sum(np.abs(H) ** 2)))
    print("  fused=%d dh=%.17g ref=%.17g hh=%.17g ref=%.17g spec_relL2=%.3g" % (fused, dh, dh_ref, hh, hh_ref,
                                                                               err_spec), flush=True)
    assert err_spec <= TOL
    assert abs(dh - dh_ref) <= TOL * hh_ref
    assert abs(hh - hh_ref) <= TOL * hh_ref
    return fused


@pytest.mark.parametrize("N", [2 ** 8, 2 ** 9, 2 ** 12, 2 ** 15, 2 ** 18, 2 ** 21, 2 ** 23])
def test_fused_fft_matches_numpy(N, monkeypatch):
    assert _case(N, 11 + N.bit_length(), monkeypatch)


@pytest.mark.parametrize("N", [64, 3 * 2 ** 12, 1000])
def test_rocfft_lengths_match_numpy(N, monkeypatch):
    assert not _case(N, 5, monkeypatch)


def test_rocfft_forced_matches_numpy(monkeypatch):
    assert not _case(2 ** 16, 9, monkeypatch, force_rocfft=True)
