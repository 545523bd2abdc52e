"""The config-1 shape over 256 seeded draws (seeds 100..355): M = 1e6, mu = 10,
a = 0.5, p0 = 8, e0 = 0.2, iota0 = 0.5, random phases and sky/spin angles,
N = 2^18 at dt = 10 s, the reference's harmonic rule (6 harmonics) and the
n = 1..30 variant.  Every waveform is compared with the oracle (the reference
CPU AAK, AAK.cc:565-591):

* plain FP64: h_I/h_II within relative L2 1e-10;
* mixed precision: noise-weighted mismatch 1 - <a|b>/sqrt(<a|a><b|b>) <= 1e-6
  under the analytic LISA PSD (instrument + 1-yr confusion, LISAnoise.py),
  and the likelihood's inner products <h|d>, <h|h> against numpy on the
  oracle's waveforms, reported; they are held to sqrt(2e-6) of
  sqrt(<h|h><d|d>), the size that a mismatch of 1e-6 alone allows
  (|<h_m - h_o|d>| <= |h_m - h_o| |d|, |h_m - h_o|^2 ~ 2 mismatch |h|^2).

The oracle runs on a pool of host threads (its ctypes calls release the GIL).
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

N, DT = 2 ** 18, 10.0
SEEDS = range(100, 356)


def _draw(seed):
    rng = np.random.default_rng(seed)
    return [0.5, 0.5, 8.0, 0.2, 1e6, 10.0, rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi),
            rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi),
            np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0]


def _weights():
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise
    f = np.fft.rfftfreq(N, DT)
    f[0] = f[1] / 10  # the reference's grid starts at f1/10 (pygpuAAK.py:72)
    return np.sqrt(DT / N) / np.sqrt(LISA_Noise(f, dur=1))


def _oracle_all(n_max):
    pars = [_draw(s) for s in SEEDS]
    threads = max(1, min(16, os.cpu_count() or 1))
    with ThreadPoolExecutor(threads) as ex:
        out = list(ex.map(lambda p: ol.o_aak_wave(p, length=N, dt=DT, n_max=n_max), pars))
    return pars, [(o["hI"], o["hII"]) for o in out]


@pytest.fixture(scope="module", params=[0, 30], ids=["rule", "n30"])
def draws(request):
    print("case config-1 draws n_max=%d: oracle on the host" % request.param, flush=True)
    pars, ref = _oracle_all(request.param)
    return request.param, pars, ref


def _inner(w, a, b):
    return sum(np.real(np.vdot(np.fft.rfft(x) * w, np.fft.rfft(y) * w)) for x, y in zip(a, b))


def test_plain_fp64_matches_the_oracle_on_every_draw(draws):
    from emri_kludge_suite_amd import GPUAAK
    n_max, pars, ref = draws
    print("case config-1 FP64 n_max=%d" % n_max, flush=True)
    g = GPUAAK(1.0, N // 10, N, 100.0, DT, True, False)
    if n_max:
        g.set_nmax(n_max)
    worst = 0.0
    for p, (rI, rII) in zip(pars, ref):
        g.gpu_gen_AAK(*p)
        _, hI, hII = g.GetWaveform()
        rel = np.sqrt((np.sum((hI[:N] - rI) ** 2) + np.sum((hII[:N] - rII) ** 2)) / (np.sum(rI ** 2) + np.sum(rII ** 2)))
        worst = max(worst, rel)
    print("FP64 worst rel L2 %.3e over %d draws" % (worst, len(pars)), flush=True)
    assert worst <= 1e-10


def test_mixed_precision_mismatch_and_inner_products(draws):
    from emri_kludge_suite_amd import GPUAAK
    n_max, pars, ref = draws
    print("case config-1 mixed n_max=%d" % n_max, flush=True)
    w = _weights()
    g = GPUAAK(1.0, N // 10, N, 100.0, DT, True, False)
    g.set_precision("mixed")
    if n_max:
        g.set_nmax(n_max)
    mism = []
    for p, (rI, rII) in zip(pars, ref):
        g.gpu_gen_AAK(*p)
        _, hI, hII = g.GetWaveform()
        a, b = (hI[:N], hII[:N]), (rI, rII)
        ab, aa, bb = _inner(w, a, b), _inner(w, a, a), _inner(w, b, b)
        mism.append(1 - ab / np.sqrt(aa * bb))
    mism = np.array(mism)
    worst = np.argsort(mism)[-3:][::-1]
    print("mixed mismatch: max %.3e, median %.3e over %d draws; worst seeds %s (%s)" % (
        mism.max(), np.median(mism), mism.size, [SEEDS[k] for k in worst],
        ", ".join("%.2e" % mism[k] for k in worst)), flush=True)
    assert mism.max() <= 1e-6
    # the likelihood batch in mixed precision against numpy on the oracle's
    # waveforms, with the first draw's oracle waveform as the data
    dI, dII = np.fft.rfft(ref[0][0]) * w, np.fft.rfft(ref[0][1]) * w
    g.input_data(dI, dII, w, w)
    like = g.LikelihoodBatch(np.array(pars))
    dd = 4 * (np.vdot(dI, dI).real + np.vdot(dII, dII).real)
    err = []
    for k, (rI, rII) in enumerate(ref):
        hI_f, hII_f = np.fft.rfft(rI) * w, np.fft.rfft(rII) * w
        hd = 4 * (np.vdot(hI_f, dI).real + np.vdot(hII_f, dII).real)
        hh = 4 * (np.vdot(hI_f, hI_f).real + np.vdot(hII_f, hII_f).real)
        scale = np.sqrt(hh * dd)
        err.append(max(abs(like[k, 0] - hd) / scale, abs(like[k, 1] - hh) / hh))
    err = np.array(err)
    print("mixed inner products: worst relative error %.3e, median %.3e" % (err.max(), np.median(err)), flush=True)
    assert err.max() <= np.sqrt(2e-6)
