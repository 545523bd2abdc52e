"""GPU: the batched likelihood caller (pyGPUAAK) on the drop-in."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KEYS = ["iota", "s", "p", "e", "ln_M", "ln_mu", "gamma", "psi", "alph", "cos_theta_S", "phi_S", "cos_theta_K",
        "phi_K", "D"]
INJ = dict(zip(KEYS, [0.5, 0.5, 9.0, 0.2, np.log(1e6), np.log(10.0), 1.0, 2.0, 3.0, 0.3, 4.0, -0.4, 5.0, 1.0]))


@pytest.mark.timeout(300)
def test_nll_vanishes_at_the_injection_and_matches_single_calls():
    from emri_kludge_suite_amd import pyGPUAAK
    print("case pygpuaak N=2^18 dt=10", flush=True)
    like = pyGPUAAK({"injection_params": INJ}, 2 ** 18, 10.0, 100.0, KEYS, backint=False, n_max=10)
    x0 = np.array([INJ[k] for k in KEYS])
    nll0 = like.getNLL(x0[:, None])
    assert abs(nll0[0]) <= 1e-10 * like.d_d
    # a small walk of templates: batched NLL equals the reference's per-template loop
    rng = np.random.default_rng(12)
    X = x0[:, None] + np.diag([0.0, 0.0, 1e-3, 1e-4, 1e-6, 0, 1e-2, 0, 0, 0, 0, 0, 0, 0]) @ rng.normal(size=(14, 6))
    X[2] = np.clip(X[2], 8.5, None)
    nll = like.getNLL(X.copy())
    assert np.all(nll >= -1e-10 * like.d_d)
    Xp = like.converter.physical_order(like.converter.convert(like.converter.recycle(X.copy())))
    for i in range(X.shape[1]):
        d_h, h_h = like.generator.WaveformThroughLikelihood(*Xp[:, i])
        assert nll[i] == pytest.approx(0.5 * (like.d_d + h_h - 2 * d_h), rel=1e-12, abs=1e-12 * like.d_d)
    snr_dh, snr_hh = like.getNLL(x0[:, None], return_snr=True)
    assert snr_dh[0] == pytest.approx(np.sqrt(like.d_d), rel=1e-10)


@pytest.mark.timeout(300)
def test_fisher_matrix_is_symmetric_positive():
    from emri_kludge_suite_amd import pyGPUAAK
    print("case fisher N=2^16 dt=10", flush=True)
    like = pyGPUAAK({"injection_params": INJ}, 2 ** 16, 10.0, 100.0, KEYS, backint=False, n_max=6, eps=1e-7)
    F = like.get_Fisher(np.array([INJ[k] for k in KEYS]))
    assert np.allclose(F, F.T)
    assert np.all(np.diag(F)[[0, 2, 3, 4, 6]] > 0)


def _oracle_spectra(like, phys):
    """Weighted spectra of the oracle's waveform at physical parameters, with
    the caller's weights (pygpuAAK.py:86-96): w rfft(h)."""
    import oracle_lib as ol
    ref = ol.o_aak_wave(phys, T_fit=like.T_fit, length=like.length, dt=like.dt, LISA=like.LISA,
                        backint=like.backint, n_max=like.n_max)
    return (like.noise_channel1_inv * np.fft.rfft(ref["hI"]), like.noise_channel2_inv * np.fft.rfft(ref["hII"]))


@pytest.mark.timeout(300)
def test_nll_and_data_norm_match_the_oracle():
    # <d|d> of the injection and NLL of nearby templates against the oracle's
    # waveforms under the same weights: 1/2 (<d|d> + <h|h> - 2 <d|h>)
    from emri_kludge_suite_amd import pyGPUAAK
    print("case pygpuaak vs oracle N=2^16 dt=10", flush=True)
    like = pyGPUAAK({"injection_params": INJ}, 2 ** 16, 10.0, 100.0, KEYS, backint=False, n_max=10)
    x0 = np.array([INJ[k] for k in KEYS])
    conv = like.converter
    phys0 = conv.physical_order(conv.convert(x0[:, None].copy()))[:, 0]
    dI, dII = _oracle_spectra(like, phys0)
    d_d = 4 * (np.vdot(dI, dI).real + np.vdot(dII, dII).real)
    print("d_d: caller %.12e oracle %.12e" % (like.d_d, d_d), flush=True)
    assert like.d_d == pytest.approx(d_d, rel=1e-9)
    rng = np.random.default_rng(21)
    X = x0[:, None] + np.diag([1e-3, 1e-3, 1e-3, 1e-3, 1e-5, 1e-3, 1e-2, 1e-2, 1e-2, 1e-2, 1e-2, 1e-2, 1e-2, 1e-2]) @ \
        rng.normal(size=(14, 4))
    nll = like.getNLL(X.copy())
    Xp = conv.physical_order(conv.convert(conv.recycle(X.copy())))
    for i in range(X.shape[1]):
        hI, hII = _oracle_spectra(like, Xp[:, i])
        h_h = 4 * (np.vdot(hI, hI).real + np.vdot(hII, hII).real)
        d_h = 4 * (np.vdot(dI, hI).real + np.vdot(dII, hII).real)
        want = 0.5 * (d_d + h_h - 2 * d_h)
        print("  template %d: NLL caller %.10e oracle %.10e" % (i, nll[i], want), flush=True)
        assert nll[i] == pytest.approx(want, rel=1e-6, abs=1e-9 * d_d)


@pytest.mark.timeout(300)
def test_fisher_matrix_matches_the_oracle():
    # central differences of the weighted spectra, <d_i h, d_j h> over both
    # channels, from the oracle's waveforms at the same 28 parameter points
    from emri_kludge_suite_amd import pyGPUAAK
    print("case fisher vs oracle N=2^15 dt=10", flush=True)
    eps = 1e-6
    like = pyGPUAAK({"injection_params": INJ}, 2 ** 15, 10.0, 100.0, KEYS, backint=False, n_max=6, eps=eps)
    x0 = np.array([INJ[k] for k in KEYS])
    F = like.get_Fisher(x0)
    conv = like.converter
    n = len(KEYS)
    x_in = np.tile(x0, (2 * n, 1))
    for i in range(n):
        x_in[2 * i, i] += eps
        x_in[2 * i + 1, i] -= eps
    Xp = conv.physical_order(conv.convert(conv.recycle(x_in.T.copy())))
    spec = [_oracle_spectra(like, Xp[:, k]) for k in range(2 * n)]
    dI = [(spec[2 * i][0] - spec[2 * i + 1][0]) / (2 * eps) for i in range(n)]
    dII = [(spec[2 * i][1] - spec[2 * i + 1][1]) / (2 * eps) for i in range(n)]
    Fo = np.array([[4 * np.real(np.vdot(dI[i], dI[j]) + np.vdot(dII[i], dII[j])) for j in range(n)]
                   for i in range(n)])
    scale = np.sqrt(np.outer(np.diag(Fo), np.diag(Fo))) + 1e-300
    err = np.abs(F - Fo) / scale
    print("fisher: max |F - F_oracle| / sqrt(F_ii F_jj) = %.3e" % err.max(), flush=True)
    assert err.max() <= 1e-4
