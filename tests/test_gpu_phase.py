"""GPU: AAK phases and frequencies (aak_gen_phases, the AAKwrapper.phase
mirror) and the AAKwrapper.wave mirror, against the oracle's restatement of
the reference's CPU AAKphase / AAKwave.  The phases come from the same exact
folds as the waveform, so they must match bit for bit; the frequencies and
eccentricity too."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _pars_dict(p, length, dt, LISA=True, backint=False):
    keys = ("iota", "s", "p", "e", "M", "mu", "gamma", "psi", "alpha", "theta_S", "phi_S", "theta_K", "phi_K", "D")
    d = dict(zip(keys, p))
    d.update(length=length, dt=dt, LISA=LISA, backint=backint, T_fit=1.0, T=1.0, f=2e-3)
    return d


@pytest.mark.parametrize("case", ["config1", "plunge-backint"])
def test_phases_bit_identical_to_oracle(case):
    from emri_kludge_suite_amd import AAKwrapper
    if case == "config1":
        p, N, dt, bi = [0.5, 0.5, 8.0, 0.2, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0], 2 ** 18, 10.0, False
    else:
        p, N, dt, bi = [np.pi / 6, 0.5, 8.0, 0.1, 1e6, 100.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0], 2 ** 20, 10.0, True
    print("case phases", case, N, dt, flush=True)
    ref = ol.o_aak_phase(p, length=N, dt=dt, backint=bi)
    out = AAKwrapper.phase(_pars_dict(p, N, dt, backint=bi))
    for name, got in zip(ol.PHASE_NAMES, out[:8]):
        assert np.array_equal(got, ref[name]), name
    assert out[8] > 0.0


def test_wave_mirror_matches_oracle():
    from emri_kludge_suite_amd import AAKwrapper
    p, N, dt = [np.pi / 6, 0.5, 8.0, 0.1, 1e6, 10.0, 0.0, 0.0, 0.0, np.pi / 3, 0.0, np.pi / 4, np.pi / 4, 5.0], 2 ** 16, 10.0
    print("case wave-mirror", flush=True)
    ref = ol.o_aak_wave(p, length=N, dt=dt, LISA=False)
    t, hI, hII, timing = AAKwrapper.wave(_pars_dict(p, N, dt, LISA=False))
    assert t.shape == (N,) and np.array_equal(t, ref["t"])
    rel = np.sqrt((np.sum((hI - ref["hI"]) ** 2) + np.sum((hII - ref["hII"]) ** 2)) /
                  (np.sum(ref["hI"] ** 2) + np.sum(ref["hII"] ** 2)))
    assert rel <= 1e-10
