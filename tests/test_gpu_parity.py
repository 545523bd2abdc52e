"""GPU parity: the HIP path against the CPU oracle on identical inputs.

Tolerances: h_I/h_II relative L2 <= 1e-10 in plain FP64; inner
product relative error <= 1e-9.  The host stage is bit-identical to the oracle
(checked on the CPU in test_host_parity.py), so the dense arrays are compared
bit for bit as well: the spline values and the phase rates must be identical,
and the phase folds must reproduce the sequential FP64 rounding exactly.
"""
import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu


def _config1(seed=1):
    rng = np.random.default_rng(seed)
    return [0.5, 0.5, 8.0, 0.2, 1e6, 10.0, rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi),
            rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi),
            np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0]


def _config2(rng):
    while True:
        pars = [0.5, rng.uniform(0, 0.95), rng.uniform(7, 15), rng.uniform(1e-3, 0.5),
                10 ** rng.uniform(5, 7), rng.uniform(5, 50), rng.uniform(0, 2 * np.pi),
                rng.uniform(0, 2 * np.pi), rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)),
                rng.uniform(0, 2 * np.pi), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi), 1.0]
        yield pars


def _rel_l2(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / max(np.sum(b ** 2), 1e-300)))


def _compare(pars, N, dt, LISA=True, backint=False, n_max=0, label=""):
    from emri_kludge_suite_amd import GPUAAK
    print("case %s N=%d dt=%g LISA=%d backint=%d n_max=%d pars=%s" % (label, N, dt, LISA, backint, n_max,
                                                                     np.array2string(np.array(pars), precision=6)), flush=True)
    ref = ol.o_aak_wave(pars, length=N, dt=dt, LISA=LISA, backint=backint, n_max=n_max, want_dense=True)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, LISA, backint)
    if n_max:
        g.set_nmax(n_max)
    g.gpu_gen_AAK(*pars)
    info = g.info()
    t, hI, hII = g.GetWaveform()
    dense = g.dense_arrays()
    out = {"info": info}
    for name in ("e", "v", "M", "S"):
        out["max_abs_" + name] = float(np.max(np.abs(dense[name] - ref["dense_" + name])))
    for name in ("gim", "Phi", "alp"):
        d = dense[name] - ref["dense_" + name]
        out["phase_mismatch_" + name] = int(np.count_nonzero(d))
        out["phase_maxabs_" + name] = float(np.max(np.abs(d)))
    out["relL2_hI"] = _rel_l2(hI[:N], ref["hI"])
    out["relL2_hII"] = _rel_l2(hII[:N], ref["hII"])
    out["t_equal"] = bool(np.array_equal(t[:N], ref["t"]))
    print(out, flush=True)
    # the parallel exact fold resolves these cases without its serial replay
    assert info["fold_fallbacks"] == 0
    assert info["nmodes"] == ref["nmodes"] and info["i_plunge"] == ref["i_plunge"]
    assert info["i_buffer"] == ref["i_buffer"] and info["K"] == ref["K"]
    return out, g, ref


def test_config1_waveform_parity():
    out, g, ref = _compare(_config1(), 2 ** 18, 10.0, label="config1")
    for name in ("e", "v", "M", "S"):
        assert out["max_abs_" + name] == 0.0, name
    assert out["t_equal"]
    for name in ("gim", "Phi", "alp"):
        assert out["phase_mismatch_" + name] == 0, name
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


def test_config1_nmax30_parity():
    out, g, ref = _compare(_config1(), 2 ** 18, 10.0, n_max=30, label="config1-n30")
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


def test_no_lisa_response_parity():
    out, g, ref = _compare(_config1(3), 2 ** 16, 10.0, LISA=False, label="ldc")
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


@pytest.mark.timeout(200)
def test_backint_plunging_parity():
    # mu=10, M=1e6, p0=8 plunges inside a year of dt=10 s samples (reference tutorial case)
    pars = [np.pi / 6, 0.5, 8.0, 0.1, 1e6, 10.0, 1.0, 2.0, 3.0, np.pi / 3, 0.0, np.pi / 4, np.pi / 4, 5.0]
    out, g, ref = _compare(pars, 2 ** 22, 8.0, backint=True, label="backint-plunge")
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


@pytest.mark.timeout(400)
def test_config2_random_parity():
    rng = np.random.default_rng(2)
    gen = _config2(rng)
    done = 0
    while done < 2:
        pars = next(gen)
        try:
            ol.o_aak_wave(pars, length=2 ** 12, dt=15.0, LISA=True)
        except RuntimeError:
            continue
        out, g, ref = _compare(pars, 2 ** 21, 15.0, n_max=50, label="config2-%d" % done)
        assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10
        done += 1


def test_likelihood_parity():
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise, generate_noise_frequencies
    N, dt = 2 ** 18, 10.0
    pars = _config1()
    ref = ol.o_aak_wave(pars, length=N, dt=dt)
    freqs = generate_noise_frequencies(N * dt, 1 / dt)
    freqs[0] = freqs[1] / 10.0
    w = 1.0 / np.sqrt(LISA_Noise(freqs, dur=1)) * np.sqrt(dt / N)
    rng = np.random.default_rng(7)
    pars2 = list(pars)
    pars2[7] += 0.3
    ref2 = ol.o_aak_wave(pars2, length=N, dt=dt)
    dI = np.fft.rfft(ref2["hI"]) * w
    dII = np.fft.rfft(ref2["hII"]) * w
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    g.input_data(dI, dII, w, w)
    like = g.WaveformThroughLikelihood(*pars)
    hI = np.fft.rfft(ref["hI"]) * w
    hII = np.fft.rfft(ref["hII"]) * w
    dh = 4 * np.sum(np.real(np.conj(hI) * dI) + np.real(np.conj(hII) * dII))
    hh = 4 * np.sum(np.abs(hI) ** 2 + np.abs(hII) ** 2)
    print("likelihood", like, dh, hh, flush=True)
    assert abs(like[0] - dh) <= 1e-9 * abs(hh)
    assert abs(like[1] - hh) <= 1e-9 * abs(hh)
    # GetWaveform after Likelihood returns the weighted spectrum
    t, sI, sII = g.GetWaveform()
    spec = sI[0::2] + 1j * sI[1::2]
    assert np.max(np.abs(spec - hI)) <= 1e-9 * np.max(np.abs(hI))


def test_serial_fold_replay_is_bit_identical():
    # the fallback the parallel fold takes when its bookkeeping overflows,
    # forced for every phase of both directions (backint)
    from emri_kludge_suite_amd import GPUAAK
    pars = [np.pi / 6, 0.5, 8.0, 0.1, 1e6, 100.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    N, dt = 2 ** 18, 10.0
    print("case serial-replay N=%d dt=%g backint=1" % (N, dt), flush=True)
    ref = ol.o_aak_wave(pars, length=N, dt=dt, backint=True, want_dense=True)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, True)
    g.force_fold_replay(True)
    g.gpu_gen_AAK(*pars)
    info = g.info()
    dense = g.dense_arrays()
    print(info, flush=True)
    assert info["fold_fallbacks"] == (6 if info["i0"] > 0 else 3)
    for name in ("gim", "Phi", "alp"):
        assert np.array_equal(dense[name], ref["dense_" + name]), name
    t, hI, hII = g.GetWaveform()
    assert _rel_l2(hI[:N], ref["hI"]) <= 1e-10 and _rel_l2(hII[:N], ref["hII"]) <= 1e-10


def test_likelihood_parity_non_power_of_two_length():
    # lengths the fused FFT does not take run rocFFT and the separate
    # weighted reduction; same bar
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.LISAnoise import LISA_Noise, generate_noise_frequencies
    N, dt = 3 * 2 ** 15, 10.0
    print("case likelihood-rocfft N=%d" % N, flush=True)
    pars = _config1(4)
    ref = ol.o_aak_wave(pars, length=N, dt=dt)
    freqs = generate_noise_frequencies(N * dt, 1 / dt)
    freqs[0] = freqs[1] / 10.0
    w = 1.0 / np.sqrt(LISA_Noise(freqs, dur=1)) * np.sqrt(dt / N)
    pars2 = list(pars)
    pars2[6] += 0.2
    ref2 = ol.o_aak_wave(pars2, length=N, dt=dt)
    dI = np.fft.rfft(ref2["hI"]) * w
    dII = np.fft.rfft(ref2["hII"]) * w
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    g.input_data(dI, dII, w, w)
    like = g.WaveformThroughLikelihood(*pars)
    batch = g.LikelihoodBatch(np.array([pars, pars2]))
    hI = np.fft.rfft(ref["hI"]) * w
    hII = np.fft.rfft(ref["hII"]) * w
    dh = 4 * np.sum(np.real(np.conj(hI) * dI) + np.real(np.conj(hII) * dII))
    hh = 4 * np.sum(np.abs(hI) ** 2 + np.abs(hII) ** 2)
    print("likelihood", like, dh, hh, flush=True)
    assert abs(like[0] - dh) <= 1e-9 * abs(hh)
    assert abs(like[1] - hh) <= 1e-9 * abs(hh)
    assert np.array_equal(batch[0], like)


def test_high_eccentricity_parity():
    # e0 = 0.7 lies above the series range: the mode sum takes the Miller path
    pars = [0.5, 0.6, 12.0, 0.7, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    out, g, ref = _compare(pars, 2 ** 16, 10.0, label="e0.7")
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


def test_mode_sum_forms_agree_across_a_plunge():
    # fast inspiral close to plunge: blocks with a wide e range fall back to the
    # per-sample Bessel evaluation (about 90 of them here), the others use the
    # block Taylor form
    pars = [0.5, 0.5818, 9.4211, 0.3929, 119708.4349, 13.4395, 4.9582, 0.2884, 0.8303, 1.5676, 4.4618, 1.3913,
            0.2321, 1.0]
    out, g, ref = _compare(pars, 2 ** 18, 15.0, n_max=50, label="fast-plunge")
    assert out["info"]["per_sample_blocks"] > 0
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10


def test_run_phase_trajectory_keeps_the_oracle_knots():
    # GPUAAK.run_phase_trajectory (AAK_manager.hh:118-132) keeps the host
    # stage in the handle; its knots and PNevolution outputs equal the free
    # function's on the oracle's sparse trajectory, and gpu_gen_AAK leaves the
    # same trajectory behind
    from emri_kludge_suite_amd import GPUAAK, PNevolution
    pars, N, dt = [0.5, 0.5, 8.0, 0.1, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0], 2 ** 18, 120.0
    print("case run_phase_trajectory N=%d dt=%g" % (N, dt), flush=True)
    r = ol.o_knots(pars, 1.0, N, dt)
    want = PNevolution(dt, N, r["par"], r["e_traj"], r["v_map"], r["M_phys"], r["M_map"], r["S_phys"],
                       r["s_map"], r["dt_map"], r["steps"], False)
    g = GPUAAK(1.0, N // 10, N, 100.0, dt, True, False)
    with pytest.raises(RuntimeError, match="no trajectory"):
        g.GetTrajectory()
    for run in (g.run_phase_trajectory, g.gpu_gen_AAK):
        run(*pars)
        got = g.GetTrajectory()
        for a, b in zip(got[:5], want[:5]):
            np.testing.assert_array_equal(a, b)
        assert got[5:] == want[5:]


def test_mixed_split_expansions_near_a_plunge():
    # mixed form, 1024-sample expansion blocks: the blocks near the plunge that
    # no order <= 8 serves are expanded again per 256 samples, so only the
    # 256-sample blocks that fail on their own are summed per sample -- the
    # same blocks as with 256-sample expansions throughout -- and the waveform
    # stays within the mixed form's mismatch bar against the oracle
    from emri_kludge_suite_amd import GPUAAK
    from emri_kludge_suite_amd.multigrid import psd_weights
    # a production-prior source plunging after 8 days: nine of its 1024-sample
    # expansion blocks need an order above 8, one 256-sample block still does
    pars = [0.5, 0.5818, 9.4211, 0.3929, 119708.4349, 13.4395, 4.9582, 0.2884, 0.8303, 1.5676, 4.4618, 1.3913,
            0.2321, 1.0]
    N, dt = 2 ** 18, 15.0
    print("case mixed split N=2^18 dt=15 n_max=50", flush=True)
    ref = ol.o_aak_wave(pars, length=N, dt=dt, n_max=50)
    w = psd_weights(N, dt, 1)
    fr = [np.fft.rfft(ref[k]) * w for k in ("hI", "hII")]
    g = GPUAAK(1.0, N // 100, N, 100.0, dt, True, False)
    g.set_nmax(50)
    g.set_precision("mixed")
    res = {}
    for xs in (0, 2):
        g.set_tuning(exp_shift=xs)
        g.gpu_gen_AAK(*pars)
        info = g.info()
        _, hI, hII = g.GetWaveform()
        fg = [np.fft.rfft(x[:N]) * w for x in (hI, hII)]
        ab = sum(np.vdot(a, b).real for a, b in zip(fr, fg))
        mm = 1 - ab / np.sqrt(sum(np.vdot(a, a).real for a in fr) * sum(np.vdot(b, b).real for b in fg))
        res[xs] = (info["per_sample_blocks"], info["split_blocks"], mm)
        print("  exp_shift %d: per-sample blocks %d, split expansion blocks %d, mismatch vs oracle %.3e"
              % (xs, *res[xs]), flush=True)
    assert res[2][1] > 0
    assert res[2][0] == res[0][0]
    assert res[0][2] <= 1e-6 and res[2][2] <= 1e-6


def test_runaway_evolution_past_the_plunge_keeps_the_series():
    # a production-prior source that plunges after 21 hours: past the plunge
    # the AK evolution runs off (e reaches 1.45 on the knots) before it is cut,
    # but no sample the mode sum evaluates lies there; the Bessel series bound
    # comes from the knots that shape the signal (e <= 0.04 here), so every
    # block keeps the block Taylor form instead of the per-sample recurrence
    pars = [0.5, 0.092934, 7.375436, 0.036833, 166902.580916, 29.647813, 1.893011, 6.231121, 3.951397, 0.509965,
            4.605533, 0.649373, 5.930568, 1.0]
    out, g, ref = _compare(pars, 2 ** 18, 15.0, n_max=50, label="runaway")
    assert out["info"]["per_sample_blocks"] == 0
    assert out["relL2_hI"] <= 1e-10 and out["relL2_hII"] <= 1e-10
