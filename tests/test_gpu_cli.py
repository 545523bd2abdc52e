"""The command-line generator AAK_Waveform (emri_kludge_suite_amd/bin) end to
end on the GPU, against the oracle: the waveform file (FP64 bar, rel L2
1e-10), the trajectory file (p, e, E, L_z, Q at the reference's downsampled
samples, rel 1e-10) and the SNRs of the info file (printed to 0.1), as
AAK_Waveform.cc:14-134 writes them."""
import math
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "emri_kludge_suite_amd", "bin", "AAK_Waveform")
MSUN_SEC = 6.67259e-11 * 1.98892e30 / 299792458.0 ** 3


def settings(tmp_path, backint, LISA, traj, SNR, timing, length, dt, p, T, f, pars):
    iota, s, _, e, M, mu, gamma, psi, alph, qS, phS, qK, phK, D = pars
    entries = ["run", *("true" if b else "false" for b in (backint, LISA, traj, SNR, timing)),
               repr(length), repr(dt), repr(p), repr(T), repr(f), "1.",
               repr(mu), repr(M), repr(s), repr(e), repr(iota), repr(gamma), repr(psi),
               repr(qS), repr(phS), repr(qK), repr(phK), repr(alph), repr(D)]
    path = tmp_path / "SetPar"
    path.write_text("# AAK_Waveform settings\n\n" + "\n\n".join(entries) + "\n")
    return str(path)


def run(path, cwd):
    print("case cli %s" % open(path).read().replace("\n\n", " ").replace("\n", " "), flush=True)
    r = subprocess.run([EXE, path], cwd=str(cwd), capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def loaded(path):
    from emri_kludge_suite_amd import _lib
    import ctypes as C
    s = _lib.SetPar()
    assert _lib.lib().aak_load_setpar(path.encode(), C.byref(s)) == 1
    return s


def snr_reference(h, dt, N):
    # InnProd (KSTools.cc:127-138) on the oracle's waveform
    from emri_kludge_suite_amd import _lib
    S = _lib.lib().aak_elisa_noise
    H = np.fft.rfft(h)
    k = np.arange(1, (N + 1) // 2)
    f = k / dt / N
    keep = (f > 1e-5) & (f < 3.0)
    Sn = np.array([S(x) for x in f[keep]])
    return 4 * np.sum(np.abs(H[k[keep]]) ** 2 / Sn) * dt / N


PARS = (0.5, 0.5, 8.0, 0.2, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0)


@pytest.mark.timeout(600)
def test_waveform_trajectory_and_snr_files(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("AAK_Waveform not built")
    N, dt = 2 ** 15, 10.0
    path = settings(tmp_path, False, True, True, True, True, N, dt, 8.0, 1.0, 2e-3, PARS)
    run(path, tmp_path)
    ref = ol.o_aak_wave(PARS, T_fit=1.0, length=N, dt=dt, LISA=True, backint=False, want_dense=True)
    w = np.loadtxt(tmp_path / "run_wave.dat")
    assert w.shape == (N, 3)
    assert np.allclose(w[:, 0], ref["t"], rtol=1e-12, atol=1e-9)
    den = np.sum(ref["hI"] ** 2) + np.sum(ref["hII"] ** 2)
    rel = math.sqrt((np.sum((w[:, 1] - ref["hI"]) ** 2) + np.sum((w[:, 2] - ref["hII"]) ** 2)) / den)
    print("cli wave: rel L2 %.3e" % rel, flush=True)
    assert rel <= 1e-10

    # trajectory rows: the reference's downsampling of p(t), e(t) up to the plunge
    iota, a, M, mu = PARS[0], PARS[1], PARS[4], PARS[5]
    ip = ref["i_plunge"]
    i_rr = int(0.001 * (MSUN_SEC * M * M / mu) / dt) + 1
    rows = [i for i in range(ip + 1) if i % i_rr == 0 or i + i_rr >= ip + 1]
    tr = np.loadtxt(tmp_path / "run_traj.dat", ndmin=2)
    assert tr.shape == (len(rows), 7)
    O = ol.oracle()
    for r, i in zip(tr, rows):
        e, v, Mi, Si = (ref["dense_" + k][i] for k in ("e", "v", "M", "S"))
        p = (1 - e * e) * (O.o_omegaphi(v, e, math.cos(iota), Si, Mi) * M * MSUN_SEC) ** (-2.0 / 3.0)
        elq = ol.o_iekg(p, e, math.cos(iota), a, freq=False)[:3]
        want = np.array([ref["t"][i], p, e, iota, *elq])
        assert np.allclose(r, want, rtol=1e-10, atol=1e-12), (i, r, want)

    # info: SNRs to the printed 0.1, and the timing lines
    info = (tmp_path / "run_info.txt").read_text().split("\n")
    vals = {ln.split(":")[0]: ln.split(":")[1].strip() for ln in info if ":" in ln}
    hI2, hII2 = snr_reference(ref["hI"], dt, N), snr_reference(ref["hII"], dt, N)
    print("cli SNR: file %s %s %s, oracle %.4f %.4f %.4f" % (vals["SNR_I"], vals["SNR_II"], vals["SNR"],
          math.sqrt(hI2), math.sqrt(hII2), math.sqrt(hI2 + hII2)), flush=True)
    for key, want in (("SNR_I", hI2), ("SNR_II", hII2), ("SNR", hI2 + hII2)):
        assert abs(float(vals[key]) - math.sqrt(want)) <= 0.05 + 1e-9 * math.sqrt(want), key
    assert float(vals["Clock ticks"]) > 0 and vals["Seconds"].endswith("s")


@pytest.mark.timeout(600)
def test_backwards_integration_derived_step_and_p(tmp_path):
    if not os.path.exists(EXE):
        pytest.fail("AAK_Waveform not built")
    N = 2 ** 15
    pars = list(PARS)
    pars[3] = 0.1
    # dt from T (0.01 of a 365-day year), p from f; the waveform ends at the plunge
    path = settings(tmp_path, True, False, False, False, True, N, -1.0, -1.0, 0.01, 4e-3, pars)
    s = loaded(path)
    pars[2] = s.p
    run(path, tmp_path)
    r2 = run(path, tmp_path)  # the second run overwrites
    assert "Output warning: Overwriting" in r2.stderr
    ref = ol.o_aak_wave(pars, T_fit=1.0, length=N, dt=s.dt, LISA=False, backint=True)
    w = np.loadtxt(tmp_path / "run_wave.dat")
    den = np.sum(ref["hI"] ** 2) + np.sum(ref["hII"] ** 2)
    rel = math.sqrt((np.sum((w[:, 1] - ref["hI"]) ** 2) + np.sum((w[:, 2] - ref["hII"]) ** 2)) / den)
    print("cli backint: dt %.6f p %.6f rel L2 %.3e" % (s.dt, s.p, rel), flush=True)
    assert rel <= 1e-10
    assert np.allclose(w[:, 0], ref["t"], rtol=1e-12, atol=1e-9)
    assert not (tmp_path / "run_traj.dat").exists()
    info = (tmp_path / "run_info.txt").read_text()
    assert "SNR" not in info and "Clock ticks:" in info
