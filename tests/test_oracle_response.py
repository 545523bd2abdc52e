"""Pins of the oracle's waveform assembly (o_waveform_from_dense; reference
AAK.cc:409-547, KSParMap.cc:532-574) by a basis-free construction.

The oracle evaluates the reference's spherical-trigonometry chain: the
orientation of L (phi_L, L.n), the angle beta and the orientation angle
2(gamma + beta), the rotation into the spin-based polarisation frame
(RotCoeff), the LISA pattern functions (cos q, phi_w, psi, F+/x of the two
Michelson channels), the Doppler phase and the sqrt(3)/2 factor.  None of that
appears below.  Instead each sample is rebuilt from vectors:

* the orbit: L = cos(lam) S + sin(lam) (cos(alpha) e1 + sin(alpha) S x e1)
  with e1 = (z - S cos(theta_K)) / sin(theta_K) (Barack & Cutler 2004, the
  precession phase alpha = 0 on the S-z plane); the pericentre P sits at the
  angle gamma from the projection of S on the orbital plane, Q = L x P;
* the harmonic-n quadrupole tensor of a Kepler ellipse,
  H_n = -a_n (PP - QQ) - b_n (PQ + QP) + c_n (PP + QQ), whose sum over n is
  the second mean-anomaly derivative of r r (the a_n, b_n, c_n are evaluated
  here with scipy's J_n), and its transverse-traceless projection about the
  source direction n, h^TT = 2 Lambda[sum_n H_n];
* the Doppler phase 2 pi f_n (R . n)/c of harmonic n at f_n = n nu + gammadot/pi,
  with R = 1 AU (cos 2 pi t/yr, sin 2 pi t/yr, 0) the detector's position;
* LISA's two channels as 90-degree Michelsons in the cartwheel frame of
  Cutler 1998: normal z' = z/2 - (sqrt 3/2) r_hat(t), arms rotating once a year
  in the retrograde sense, channel II turned by 45 degrees;
  h_I = (sqrt 3/2) D_I : h^TT with D = (x'x' - y'y')/2.

A sign slip or a swapped angle anywhere in the trigonometric chain fails this
at some of the 120 random orientations; the two agree to ~1e-14.
"""
import numpy as np
import pytest
from scipy.special import jv

import oracle_lib as ol

O = ol.oracle()
YEAR = 31558149.8                      # Constants.h
AUSEC = 1.4959787066e11 / 299792458.0  # Constants.h: AU / c
MSUN_SEC = 6.67259e-11 * 1.98892e30 / 299792458.0 ** 3
Z = np.array([0.0, 0.0, 1.0])


def _unit(theta, phi):
    return np.array([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])


def _abc(n, e, amp, phase):
    J = jv(np.arange(n - 2, n + 3), n * e)
    a = -n * amp * (J[0] - 2 * e * J[1] + 2 / n * J[2] + 2 * e * J[3] - J[4]) * np.cos(phase)
    b = -n * amp * np.sqrt(1 - e * e) * (J[0] - 2 * J[2] + J[4]) * np.sin(phase)
    c = 2 * amp * J[2] * np.cos(phase)
    return a, b, c


def _h_tt(sample, lam, qS, phS, qK, phK, nmodes, zeta, Mphys, doppler):
    t, e, v, M, S_, gim, Phi, alp, nu, gimdot = sample
    n = _unit(qS, phS)
    S = _unit(qK, phK)
    e1 = (Z - S * np.cos(qK)) / np.sin(qK)
    L = np.cos(lam) * S + np.sin(lam) * (np.cos(alp) * e1 + np.sin(alp) * np.cross(S, e1))
    Sperp = S - L * np.dot(L, S)
    Sperp /= np.linalg.norm(Sperp)
    P = np.cos(gim) * Sperp + np.sin(gim) * np.cross(L, Sperp)
    Q = np.cross(L, P)
    amp = (O.o_omegaphi(v, e, np.cos(lam), S_, M) * Mphys * MSUN_SEC) ** (2 / 3) * zeta
    orb = 2 * np.pi * t / YEAR
    R = AUSEC * np.array([np.cos(orb), np.sin(orb), 0.0])
    PP, QQ, PQ = np.outer(P, P), np.outer(Q, Q), np.outer(P, Q) + np.outer(Q, P)
    H = np.zeros((3, 3))
    for k in range(1, nmodes + 1):
        phase = k * Phi + (2 * np.pi * (k * nu + gimdot / np.pi) * np.dot(R, n) if doppler else 0.0)
        a, b, c = _abc(k, e, amp, phase)
        H += -a * (PP - QQ) - b * PQ + c * (PP + QQ)
    Pt = np.eye(3) - np.outer(n, n)
    return 2 * (Pt @ H @ Pt - 0.5 * Pt * np.trace(Pt @ H)), n, orb


def _lisa_channels(h, orb):
    er = np.array([np.cos(orb), np.sin(orb), 0.0])
    ephi = np.array([-np.sin(orb), np.cos(orb), 0.0])
    # in-plane pair (v1, v2, z') right-handed, then the arms turned by -orb about z'
    v1, v2 = -ephi, np.sqrt(3) / 2 * Z + 0.5 * er
    x = np.cos(orb) * v1 - np.sin(orb) * v2
    y = np.sin(orb) * v1 + np.cos(orb) * v2
    x2, y2 = (x + y) / np.sqrt(2), (y - x) / np.sqrt(2)
    DI = 0.5 * (np.outer(x, x) - np.outer(y, y))
    DII = 0.5 * (np.outer(x2, x2) - np.outer(y2, y2))
    return np.sqrt(3) / 2 * np.sum(DI * h), np.sqrt(3) / 2 * np.sum(DII * h)


def _random_case(rng, L):
    ang = (rng.uniform(0.05, np.pi - 0.05), np.arccos(rng.uniform(-1, 1)), rng.uniform(0, 2 * np.pi),
           np.arccos(rng.uniform(-0.99, 0.99)), rng.uniform(0, 2 * np.pi))
    dense = np.zeros((10, L + 1))
    dense[0] = np.sort(rng.uniform(0, 3e7, L + 1))
    dense[1] = rng.uniform(0.01, 0.6)
    dense[2] = rng.uniform(0.15, 0.4)
    dense[3] = 1e6 * MSUN_SEC * rng.uniform(0.99, 1.01)
    dense[4] = rng.uniform(0.0, 0.95)
    dense[5] = rng.uniform(0, 60, L + 1)
    dense[6] = rng.uniform(0, 1e4, L + 1)
    dense[7] = rng.uniform(0, 60, L + 1)
    dense[8] = rng.uniform(1e-5, 2e-3, L + 1)
    dense[9] = rng.uniform(1e-7, 1e-5, L + 1)
    return ang, dense


def _oracle_waveform(dense, L, nmodes, zeta, Mphys, ang, mich):
    t, hI, hII = np.zeros(L), np.zeros(L), np.zeros(L)
    flat = np.ascontiguousarray(dense.ravel())
    lam, qS, phS, qK, phK = ang
    # i_plunge = L - 1: no taper on the samples compared
    rc = O.o_waveform_from_dense(ol._ptr(flat), L, 10.0, nmodes, zeta, Mphys, lam, qS, phS, qK, phK, int(mich),
                                 L - 1, 10, ol._ptr(t), ol._ptr(hI), ol._ptr(hII))
    assert rc == 0
    return hI, hII


@pytest.mark.parametrize("seed", [31, 32, 33, 34])
def test_lisa_response_is_the_detector_tensor_on_the_tt_quadrupole(seed):
    rng = np.random.default_rng(seed)
    L, nmodes, zeta, Mphys = 4, 14, 1e-21, 1e6
    worst = 0.0
    for _ in range(30):
        ang, dense = _random_case(rng, L)
        hI, hII = _oracle_waveform(dense, L, nmodes, zeta, Mphys, ang, True)
        for i in range(L - 1):
            h, _, orb = _h_tt(dense[:, i], *ang, nmodes, zeta, Mphys, doppler=True)
            eI, eII = _lisa_channels(h, orb)
            scale = np.sqrt(0.75 * np.sum(h * h))  # |h^TT| sets the size of either channel
            worst = max(worst, abs(eI - hI[i]) / scale, abs(eII - hII[i]) / scale)
    assert worst < 1e-12


def test_ldc_response_is_the_tt_quadrupole_in_the_sky_basis():
    # LISA = false (AAK.cc:462-479): the two outputs are the + and x
    # polarisations in the basis p = d n / d theta_S, q = p x n, without Doppler
    rng = np.random.default_rng(35)
    L, nmodes, zeta, Mphys = 4, 10, 1e-21, 1e6
    worst = 0.0
    for _ in range(40):
        ang, dense = _random_case(rng, L)
        hI, hII = _oracle_waveform(dense, L, nmodes, zeta, Mphys, ang, False)
        lam, qS, phS, qK, phK = ang
        p = np.array([np.cos(qS) * np.cos(phS), np.cos(qS) * np.sin(phS), -np.sin(qS)])
        for i in range(L - 1):
            h, n, _ = _h_tt(dense[:, i], *ang, nmodes, zeta, Mphys, doppler=False)
            q = np.cross(p, n)
            plus, cross = 0.5 * (p @ h @ p - q @ h @ q), p @ h @ q
            scale = np.sqrt(0.5 * np.sum(h * h))
            worst = max(worst, abs(plus - hI[i]) / scale, abs(cross - hII[i]) / scale)
    assert worst < 1e-12


def test_channels_are_invariant_under_the_polarisation_frame():
    # the LDC channels are the two polarisations in some sky basis, so their
    # quadrature sum h+^2 + hx^2 = |h^TT|^2 / 2 whatever that basis is
    rng = np.random.default_rng(36)
    L = 4
    for _ in range(20):
        ang, dense = _random_case(rng, L)
        hI, hII = _oracle_waveform(dense, L, 8, 1e-21, 1e6, ang, False)
        for i in range(L - 1):
            h, _, _ = _h_tt(dense[:, i], *ang, 8, 1e-21, 1e6, doppler=False)
            assert hI[i] ** 2 + hII[i] ** 2 == pytest.approx(0.5 * np.sum(h * h), rel=1e-12)


def test_planck_taper_window():
    # AAK.cc:536-547: past i_plunge the waveform is multiplied by a Planck-taper
    # window w(t) that falls from 1 at t_plunge to 0 at t_zero = t_plunge +
    # i_buffer dt.  Measured as the ratio of a tapered to an untapered run:
    # w(mid + u) + w(mid - u) = 1 (the exponent is odd about the midpoint),
    # w decreases, w -> 1 and 0 at the ends, and the waveform is 0 from t_zero.
    rng = np.random.default_rng(37)
    L, ip, ib, dt = 400, 100, 200, 10.0
    ang, dense = _random_case(rng, L)
    dense[0] = np.arange(L + 1) * dt
    flat = np.ascontiguousarray(dense.ravel())
    out = []
    for i_plunge in (L - 1, ip):
        t, hI, hII = np.zeros(L), np.zeros(L), np.zeros(L)
        O.o_waveform_from_dense(ol._ptr(flat), L, dt, 6, 1e-21, 1e6, *ang, 1, i_plunge, ib,
                                ol._ptr(t), ol._ptr(hI), ol._ptr(hII))
        out.append((hI.copy(), hII.copy()))
    (uI, uII), (tI, tII) = out
    np.testing.assert_array_equal(tI[:ip + 1], uI[:ip + 1])
    w = tI[ip + 1:ip + ib] / uI[ip + 1:ip + ib]
    np.testing.assert_allclose(tII[ip + 1:ip + ib] / uII[ip + 1:ip + ib], w, rtol=1e-12)
    assert np.all(np.diff(w) <= 1e-15)
    assert w[0] > 1 - 1e-12 and w[-1] < 1e-12
    # sample ip + k sits at u = k - ib/2 from the midpoint
    k = np.arange(1, ib)
    sym = {kk: ww for kk, ww in zip(k, w)}
    for kk in range(1, ib // 2):
        assert sym[ib // 2 + kk] + sym[ib // 2 - kk] == pytest.approx(1.0, abs=1e-12)
    assert sym[ib // 2] == pytest.approx(0.5, abs=1e-14)
    assert np.all(tI[ip + ib:] == 0) and np.all(tII[ip + ib:] == 0)
