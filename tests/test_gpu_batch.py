"""GPU: the batched likelihood path and the Phidot guard.

aak_likelihood_batch runs the same kernels as gpu_gen_AAK + Likelihood, with
the host stages on a thread pool and the knot-table uploads on a second
stream; every row must equal the single-call result bit for bit, templates
whose host stage fails must come back as NaN rows with their status, and
repeated batches must agree.  The guard case checks a template with
non-positive Phidot samples against the oracle.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

N2, DT2 = 2 ** 20, 15.0


def _prior(rng, n):
    out = []
    for _ in range(n):
        out.append([0.5, rng.uniform(0, 0.95), rng.uniform(7, 15), rng.uniform(1e-3, 0.5),
                    10 ** rng.uniform(5, 7), rng.uniform(5, 50), rng.uniform(0, 2 * np.pi),
                    rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)),
                    rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0])
    return np.array(out)


def _handle_with_data(N, dt, n_max, seed=2):
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise, generate_noise_frequencies
    freqs = generate_noise_frequencies(N * dt, 1 / dt)
    freqs[0] = freqs[1] / 10.0
    w = 1.0 / np.sqrt(LISA_Noise(freqs, dur=1)) * np.sqrt(dt / N)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    g.set_nmax(n_max)
    rng = np.random.default_rng(seed)
    while True:
        inj = _prior(rng, 1)[0]
        try:
            g.gpu_gen_AAK(*inj)
            break
        except RuntimeError:
            continue
    _, hI, hII = g.GetWaveform()
    g.input_data(np.fft.rfft(hI[:N]) * w, np.fft.rfft(hII[:N]) * w, w, w)
    return g


@pytest.mark.timeout(300)
def test_batch_equals_single_calls():
    print("case batch-vs-single N=2^20 dt=15 n_max=20 seeds 2/3", flush=True)
    g = _handle_with_data(N2, DT2, 20)
    P = _prior(np.random.default_rng(3), 10)
    P[4, 2] = 3.0  # p0 inside the LSO: host stage fails
    like, st = g.LikelihoodBatch(P, return_status=True)
    print("status", st.tolist(), flush=True)
    assert st[4] < 0 and np.all(np.isnan(like[4]))
    plunged = 0
    for i in range(len(P)):
        if st[i] < 0:
            with pytest.raises(RuntimeError):
                g.WaveformThroughLikelihood(*P[i])
            continue
        single = g.WaveformThroughLikelihood(*P[i])
        inf = g.info()
        # batches compute rates only up to the last folded sample; single
        # calls compute them everywhere
        plunged += inf["i_plunge"] + inf["i_buffer"] < N2 - 1
        assert np.array_equal(like[i], single), (i, like[i], single)
    print("plunging templates", plunged, flush=True)
    assert plunged >= 1


@pytest.mark.timeout(300)
def test_batch_with_trimmed_tables_equals_single_calls():
    # sources that plunge early into clamp cycles: the batch uploads their
    # tables only up to the last sample it evaluates, single calls upload
    # them whole (every sample is evaluated there); the rows must agree
    from test_host_parity import CLAMP_CYCLE_SOURCES
    N, dt = 2 ** 21, 15.0
    print("case batch-trimmed-tables N=2^21 dt=15 n_max=20", flush=True)
    g = _handle_with_data(N, dt, 20, seed=6)
    P = np.array(CLAMP_CYCLE_SOURCES + [list(_prior(np.random.default_rng(7), 1)[0])])
    like, st = g.LikelihoodBatch(P, return_status=True)
    assert np.all(st[:3] == 0), st
    for i in range(len(P)):
        if st[i] < 0:
            continue
        single = g.WaveformThroughLikelihood(*P[i])
        if i < 3:
            inf = g.info()
            assert inf["i_plunge"] + inf["i_buffer"] < N // 2, inf
        assert np.array_equal(like[i], single), (i, like[i], single)


@pytest.mark.timeout(300)
def test_batch_repeatable_and_thread_count_independent():
    print("case batch-repeat N=2^20 dt=15 n_max=12 seed 5", flush=True)
    g = _handle_with_data(N2, DT2, 12, seed=4)
    P = _prior(np.random.default_rng(5), 24)
    a = g.LikelihoodBatch(P)
    g.set_host_threads(1)
    b = g.LikelihoodBatch(P)
    g.set_host_threads(3)
    c = g.LikelihoodBatch(P[::-1])[::-1]
    ok = ~np.isnan(a[:, 0])
    assert ok.sum() >= 12
    assert np.array_equal(a[ok], b[ok]) and np.array_equal(a[ok], c[ok])
    assert np.array_equal(np.isnan(a), np.isnan(b))


@pytest.mark.timeout(300)
def test_batch_lane_count_independent(monkeypatch):
    # batches spread templates over AAK_LANES concurrent streams (read when a
    # handle is created); the results must not depend on the lane count
    print("case batch-lanes N=2^20 dt=15 n_max=12 seed 7, lanes 1/2/3", flush=True)
    P = _prior(np.random.default_rng(7), 16)
    out = {}
    for lanes in (1, 2, 3):
        monkeypatch.setenv("AAK_LANES", str(lanes))
        g = _handle_with_data(N2, DT2, 12, seed=4)
        out[lanes] = g.LikelihoodBatch(P)
        del g
    ok = ~np.isnan(out[1][:, 0])
    assert ok.sum() >= 8
    for lanes in (2, 3):
        assert np.array_equal(np.isnan(out[1]), np.isnan(out[lanes]))
        assert np.array_equal(out[1][ok], out[lanes][ok]), lanes


def test_batch_profiled_stage_times():
    print("case batch-profiling N=2^18 dt=10", flush=True)
    g = _handle_with_data(2 ** 18, 10.0, 8)
    P = _prior(np.random.default_rng(6), 6)
    g.set_profiling(True)
    g.LikelihoodBatch(P)
    tm = g.batch_timings()
    print(tm, flush=True)
    assert tm["modesum"] > 0 and tm["rates"] > 0 and tm["fft"] > 0 and tm["dot"] > 0


@pytest.mark.timeout(400)
def test_phidot_guard_against_oracle():
    # find a template of the production prior whose dense Phidot has non-positive
    # samples in the folded range, and check the whole waveform against the oracle
    from emri_kludge_suite_amd import GPUAAK
    N, dt = 2 ** 21, 15.0
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    g.set_nmax(8)
    rng = np.random.default_rng(3)
    found = None
    for k in range(40):
        pars = _prior(rng, 1)[0]
        try:
            g.gpu_gen_AAK(*pars)
        except RuntimeError:
            continue
        if g.info()["nonpositive_phidot"] > 0:
            found = pars
            break
    if found is None:
        pytest.skip("no template with non-positive Phidot among the first 40 draws")
    print("case guard pars=%s info=%s" % (np.array2string(found, precision=6), g.info()), flush=True)
    ref = ol.o_aak_wave(list(found), length=N, dt=dt, LISA=True, n_max=8, want_dense=True)
    dense = g.dense_arrays()
    for name in ("gim", "Phi", "alp"):
        assert np.array_equal(dense[name], ref["dense_" + name]), name
    _, hI, hII = g.GetWaveform()
    rel = np.sqrt((np.sum((hI[:N] - ref["hI"]) ** 2) + np.sum((hII[:N] - ref["hII"]) ** 2)) /
                  (np.sum(ref["hI"] ** 2) + np.sum(ref["hII"] ** 2)))
    assert rel <= 1e-10


def _weights(N, dt):
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise, generate_noise_frequencies
    freqs = generate_noise_frequencies(N * dt, 1 / dt)
    freqs[0] = freqs[1] / 10.0
    return 1.0 / np.sqrt(LISA_Noise(freqs, dur=1)) * np.sqrt(dt / N)


def _mismatch(a, b, w, N):
    fa = [np.fft.rfft(x[:N]) * w for x in a]
    fb = [np.fft.rfft(x[:N]) * w for x in b]
    ab = sum(np.real(np.vdot(x, y)) for x, y in zip(fa, fb))
    aa = sum(np.real(np.vdot(x, x)) for x in fa)
    bb = sum(np.real(np.vdot(y, y)) for y in fb)
    return 1.0 - ab / np.sqrt(aa * bb)


@pytest.mark.timeout(300)
def test_mixed_precision_mismatch_below_1e6():
    # mixed precision against the FP64 path on production-prior templates at
    # N=2^21, dt=15 s, n=1..50: noise-weighted mismatch <= 1e-6
    from emri_kludge_suite_amd import GPUAAK
    N, dt = 2 ** 21, 15.0
    print("case mixed-mismatch N=2^21 dt=15 n_max=50 seed 8", flush=True)
    w = _weights(N, dt)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    g.set_nmax(50)
    rng = np.random.default_rng(8)
    done, worst = 0, 0.0
    while done < 6:
        pars = _prior(rng, 1)[0]
        try:
            g.set_precision("fp64")
            g.gpu_gen_AAK(*pars)
        except RuntimeError:
            continue
        _, aI, aII = g.GetWaveform()
        g.set_precision("mixed")
        g.gpu_gen_AAK(*pars)
        _, bI, bII = g.GetWaveform()
        mm = _mismatch((aI, aII), (bI, bII), w, N)
        # the bench's device-side form of the same mismatch: the FP64
        # template's weighted spectrum as the data of the mixed template
        dI, dII = np.fft.rfft(aI[:N]) * w, np.fft.rfft(aII[:N]) * w
        g.input_data(dI, dII, w, w)
        g.set_precision("fp64")
        g.gpu_gen_AAK(*pars)
        hh_64 = g.Likelihood()[1]
        _, sI, sII = g.GetWaveform()
        np.testing.assert_allclose(sI.view(np.complex128), dI, rtol=0, atol=1e-12 * np.max(np.abs(dI)))
        g.input_data(sI.view(np.complex128), sII.view(np.complex128), w, w)
        g.set_precision("mixed")
        g.gpu_gen_AAK(*pars)
        cross, hh_mx = g.Likelihood()
        mm_dev = 1.0 - cross / np.sqrt(hh_mx * hh_64)
        print("template", done, "mismatch", mm, "device form", mm_dev, flush=True)
        assert abs(mm_dev - mm) <= 1e-12
        worst = max(worst, mm)
        done += 1
    assert worst <= 1e-6


@pytest.mark.timeout(300)
def test_mixed_batch_equals_mixed_single_calls():
    print("case mixed-batch N=2^20 dt=15 n_max=20", flush=True)
    g = _handle_with_data(N2, DT2, 20)
    g.set_precision("mixed")
    P = _prior(np.random.default_rng(9), 8)
    like, st = g.LikelihoodBatch(P, return_status=True)
    g.set_precision("fp64")
    ref = g.LikelihoodBatch(P)
    g.set_precision("mixed")
    for i in range(len(P)):
        if st[i] < 0:
            continue
        assert np.array_equal(like[i], g.WaveformThroughLikelihood(*P[i]))
        # the likelihood terms agree with the FP64 path to well within the mismatch budget
        assert abs(like[i][1] - ref[i][1]) <= 1e-5 * abs(ref[i][1])


@pytest.mark.timeout(300)
def test_batch_edge_cases():
    # no data yet, an empty batch, a batch whose every host stage fails, and
    # a one-template batch on the same handle afterwards
    print("case batch-edges N=2^18 dt=10 n_max=10 seeds 5/6", flush=True)
    from emri_kludge_suite_amd import GPUAAK
    N, dt = 2 ** 18, 10.0
    g0 = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    with pytest.raises(RuntimeError):
        g0.LikelihoodBatch(_prior(np.random.default_rng(5), 1))
    g = _handle_with_data(N, dt, 10)
    assert g.LikelihoodBatch(np.zeros((0, 14))).shape == (0, 2)
    P = _prior(np.random.default_rng(6), 3)
    bad = P.copy()
    bad[:, 2] = 3.0  # p0 inside the LSO
    like, st = g.LikelihoodBatch(bad, return_status=True)
    assert np.all(st < 0) and np.all(np.isnan(like))
    rng = np.random.default_rng(7)
    while True:
        tpl = _prior(rng, 1)
        try:
            single = g.WaveformThroughLikelihood(*tpl[0])
            break
        except RuntimeError:
            continue
    one = g.LikelihoodBatch(tpl)
    assert np.array_equal(one[0], single)
