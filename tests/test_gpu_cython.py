"""GPU parity through the Cython binding (gpuAAK.pyx on the C ABI): the same
waveform and likelihood as the ctypes mirror, bit for bit, and the oracle's
waveform within the FP64 bar (relative L2 <= 1e-10)."""
import numpy as np
import pytest

import oracle_lib as ol
from test_cython_binding import load_gpuAAK

pytestmark = pytest.mark.gpu


def test_cython_module_matches_the_mirror_and_the_oracle():
    m = load_gpuAAK()
    from emri_kludge_suite_amd import GPUAAK
    N, dt = 2 ** 16, 10.0
    pars = [0.5, 0.5, 8.0, 0.2, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    print("case cython N=%d dt=%g" % (N, dt), flush=True)
    ref = ol.o_aak_wave(pars, length=N, dt=dt)
    nb = N // 2 + 1
    w = np.full(nb, 1e20)
    dI, dII = np.fft.rfft(ref["hI"]) * w, np.fft.rfft(ref["hII"]) * w
    out = []
    for G in (m.GPUAAK, GPUAAK):
        g = G(1.0, N // 10, N, 100.0, dt, True, False)
        g.input_data(dI, dII, w, w)
        g.gpu_gen_AAK(*pars)
        t, hI, hII = g.GetWaveform()
        like = g.Likelihood()
        batch = g.LikelihoodBatch(np.array([pars, pars]))
        out.append((hI.copy(), hII.copy(), np.array(like), batch))
    (cI, cII, cl, cb), (pI, pII, pl, pb) = out
    np.testing.assert_array_equal(cI, pI)
    np.testing.assert_array_equal(cII, pII)
    np.testing.assert_array_equal(cl, pl)
    np.testing.assert_array_equal(cb, pb)
    np.testing.assert_array_equal(cb[0], cl)
    rel = np.sqrt((np.sum((cI[:N] - ref["hI"]) ** 2) + np.sum((cII[:N] - ref["hII"]) ** 2)) /
                  (np.sum(ref["hI"] ** 2) + np.sum(ref["hII"] ** 2)))
    assert rel <= 1e-10
