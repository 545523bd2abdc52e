"""Batches of uneven duration, sampling step and harmonic count in one call
(emri_kludge_suite_amd/multigrid.py): every row equals, bit for bit, a
single-template call on a handle of its own grid, and three rows match the
oracle's waveforms within the FP64 bar (inner products to 1e-9)."""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _workload(n, seed):
    from emri_kludge_suite_amd.multigrid import config5_workload
    P, _, _, nmax, _ = config5_workload(n, seed)
    rng = np.random.default_rng(seed + 100)
    # short grids keep the test quick: N in {2^15, 2^16}, dt in {10, 15}
    N = rng.choice([2 ** 15, 2 ** 16], n)
    dt = rng.choice([10.0, 15.0], n)
    return P, N, dt, nmax


@pytest.mark.timeout(600)
def test_rows_equal_single_calls_on_their_grid():
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.multigrid import MultiGridBatch, injected_data
    P, N, dt, nmax = _workload(24, 61)
    print("case multigrid rows=%d grids=%s" % (len(P), sorted(set(zip(N.tolist(), dt.tolist())))), flush=True)
    mg = MultiGridBatch(precision="mixed", tuning={})
    from emri_kludge_suite_amd.multigrid import injection_source
    inj = injection_source(7)
    data = {}
    for k in sorted(set(zip(N.tolist(), dt.tolist()))):
        data[k] = injected_data(mg, inj, k[0], k[1], seed=7)
    like, st = mg.likelihood(P, N, dt, nmax)
    for (Ng, dtg), (d, w) in data.items():
        g = GPUAAK(1.0, Ng // 100, Ng, 100.0, dtg, True, False)
        g.set_precision("mixed")
        g.input_data(d[0], d[1], w, w)
        for i in np.nonzero((N == Ng) & (dt == dtg))[0]:
            if st[i] < 0:
                continue
            g.set_nmax(int(nmax[i]))
            single = g.WaveformThroughLikelihood(*P[i])
            assert np.array_equal(like[i], single), (i, like[i], single)
    assert np.count_nonzero(st == 0) >= len(P) // 2


@pytest.mark.timeout(600)
def test_three_rows_match_the_oracle():
    from emri_kludge_suite_amd.multigrid import MultiGridBatch, psd_weights
    P, N, dt, nmax = _workload(12, 62)
    mg = MultiGridBatch(precision="fp64", tuning={})
    checked = 0
    for i in range(len(P)):
        Ng, dtg = int(N[i]), float(dt[i])
        try:
            ref = ol.o_aak_wave(P[i], length=Ng, dt=dtg, n_max=int(nmax[i]))
        except RuntimeError:
            continue
        print("case multigrid oracle row %d N=%d dt=%g nmax=%d" % (i, Ng, dtg, nmax[i]), flush=True)
        w = psd_weights(Ng, dtg, 1)
        dI, dII = np.fft.rfft(ref["hI"]) * w, np.fft.rfft(ref["hII"]) * w
        mg.input_data(Ng, dtg, dI, dII, w, w)
        like, st = mg.likelihood(P[i:i + 1], [Ng], [dtg], nmax[i:i + 1])
        assert st[0] == 0
        hh = 4 * (np.vdot(dI, dI).real + np.vdot(dII, dII).real)
        assert like[0, 1] == pytest.approx(hh, rel=1e-9)
        assert like[0, 0] == pytest.approx(hh, rel=1e-9)
        checked += 1
        if checked == 3:
            break
    assert checked == 3
