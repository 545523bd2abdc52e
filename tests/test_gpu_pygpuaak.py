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
