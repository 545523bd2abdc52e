"""Pins of the CPU oracle by mathematics and by independent libraries.

Each check is chosen so that a plausible transcription mistake (a dropped
term, a wrong sign, an index off by one) fails it; none of them retypes the
oracle's formula.  Library pins: scipy.special (Carlson integrals, Legendre
forms, J_n), scipy.interpolate.CubicSpline (natural spline), numpy lstsq
(anchored quadratic fit), numpy.roots (quartic), scipy.optimize.root with
MINPACK hybr (frequency map).  Closed forms: Schwarzschild and Kerr
circular-equatorial geodesics, Kepler's equation, Keplerian quadrupole
(Kapteyn series), Newtonian and 1PN limits of the AK rates.
"""
import numpy as np
import pytest
from scipy import special
from scipy.integrate import quad
from scipy.interpolate import CubicSpline
from scipy.optimize import brentq, root

import oracle_lib as ol

O = ol.oracle()


# ------------------------------------------------------------ elliptic integrals
def test_carlson_against_scipy():
    rng = np.random.default_rng(21)
    for x, y, z, p in rng.uniform(1e-3, 30.0, size=(400, 4)):
        assert O.o_rf(x, y, z) == pytest.approx(special.elliprf(x, y, z), rel=2e-15)
        assert O.o_rd(x, y, z) == pytest.approx(special.elliprd(x, y, z), rel=2e-15)
        assert O.o_rj(x, y, z, p) == pytest.approx(special.elliprj(x, y, z, p), rel=4e-15)
        assert O.o_rc(x, y) == pytest.approx(special.elliprc(x, y), rel=2e-15)


def test_legendre_forms_against_scipy_and_quadrature():
    rng = np.random.default_rng(22)
    for phi, n, k in zip(rng.uniform(0, np.pi / 2, 200), rng.uniform(-0.95, 0.95, 200), rng.uniform(0, 0.99, 200)):
        assert O.o_ellf(phi, k) == pytest.approx(special.ellipkinc(phi, k * k), rel=4e-15)
        assert O.o_elle(phi, k) == pytest.approx(special.ellipeinc(phi, k * k), rel=4e-15)
        # third kind in the Numerical-Recipes sign convention: 1 / (1 + n sin^2)
        ref = quad(lambda t: 1.0 / ((1 + n * np.sin(t) ** 2) * np.sqrt(1 - (k * np.sin(t)) ** 2)), 0, phi,
                   epsabs=1e-16, epsrel=1e-13)[0]
        assert O.o_ellpi(phi, n, k) == pytest.approx(ref, rel=1e-12)


# ------------------------------------------------------------------ Bessel J_n
def test_bessel_against_scipy():
    out = np.zeros(90)
    for x in (1e-4, 0.3, 1.0, 5.5, 17.0, 33.0, 47.5):
        assert O.o_bessel_jn_array(0, 89, x, ol._ptr(out)) == 0
        ref = special.jv(np.arange(90), x)
        np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-16)


# ------------------------------------------------------- geodesic closed forms
def test_schwarzschild_energy_and_angular_momentum():
    # a -> 0: E^2 = ((p-2)^2 - 4e^2) / (p (p-3-e^2)), L^2 = p^2/(p-3-e^2) = Lz^2 + Q
    for p, e, cosi in ((10.0, 0.3, 0.6), (8.0, 0.5, -0.3), (12.0, 0.01, 0.95), (7.5, 0.6, 0.2)):
        g = ol.o_iekg(p, e, cosi, 0.0, freq=False)
        L2 = p * p / (p - 3 - e * e)
        assert g[0] ** 2 == pytest.approx(((p - 2) ** 2 - 4 * e * e) / (p * (p - 3 - e * e)), rel=1e-13)
        assert g[1] ** 2 + g[2] == pytest.approx(L2, rel=1e-11)
        assert g[1] == pytest.approx(cosi * np.sqrt(L2), rel=1e-12)


def test_schwarzschild_limit_of_frequencies():
    # a -> 0: Omega_theta = |Omega_phi| and Omega_phi/Omega_r = 2 sqrt(p/(p-6+2e)) K(4e/(p-6+2e)) / pi;
    # the deviation is O(a) (the reference formula divides by a, so a = 0 itself is singular)
    for p, e, cosi in ((10.0, 0.3, 0.6), (8.0, 0.5, -0.3)):
        ratio = 2 * np.sqrt(p / (p - 6 + 2 * e)) * special.ellipk(4 * e / (p - 6 + 2 * e)) / np.pi
        for a in (1e-2, 1e-3):
            Om = ol.o_iekg(p, e, cosi, a)[13:16]
            assert abs(Om[1] / abs(Om[2]) - 1) < 0.2 * a
            assert abs(abs(Om[2]) / Om[0] / ratio - 1) < 0.5 * a
            assert np.sign(Om[2]) == np.sign(cosi)


def test_kerr_circular_equatorial_closed_forms():
    # Bardeen-Press-Teukolsky: E, Lz and the three fundamental frequencies (cycles per M)
    for r, a in ((8.0, 0.5), (12.0, 0.9), (6.5, 0.3)):
        g = ol.o_iekg(r, 1e-6, 1.0, a)
        x = r ** -1.5
        den = np.sqrt(1 - 3 / r + 2 * a * x)
        assert g[0] == pytest.approx((1 - 2 / r + a * x) / den, rel=1e-9)
        assert g[1] == pytest.approx(np.sqrt(r) * (1 - 2 * a * x + a * a / r / r) / den, rel=1e-9)
        Op = 1 / (r ** 1.5 + a)
        Or = Op * np.sqrt(1 - 6 / r + 8 * a * x - 3 * a * a / r / r)
        Ot = Op * np.sqrt(1 - 4 * a * x + 3 * a * a / r / r)
        np.testing.assert_allclose(g[13:16] * 2 * np.pi, [Or, Ot, Op], rtol=1e-8)


def test_geodesic_turning_points_are_roots_of_the_radial_potential():
    # R(r) = [E(r^2+a^2) - a Lz]^2 - Delta [r^2 + (Lz - a E)^2 + Q] vanishes at r_p and r_a
    for p, e, cosi, a in ((8.0, 0.2, np.cos(0.5), 0.5), (11.0, 0.6, -0.4, 0.9), (7.0, 0.1, 0.3, 0.2)):
        E, Lz, Q = ol.o_iekg(p, e, cosi, a)[:3]
        for r in (p / (1 + e), p / (1 - e)):
            R = (E * (r * r + a * a) - a * Lz) ** 2 - (r * r - 2 * r + a * a) * (r * r + (Lz - a * E) ** 2 + Q)
            assert abs(R) < 1e-10 * r ** 4


def test_quartic_roots_against_numpy():
    rng = np.random.default_rng(23)
    for _ in range(200):
        c = rng.normal(size=5)
        re, im = np.zeros(4), np.zeros(4)
        O.o_quartic_roots(ol._ptr(c), ol._ptr(re), ol._ptr(im))
        ref = np.roots(c[::-1])
        got = re + 1j * im
        for z in ref:
            assert np.min(np.abs(got - z)) < 1e-9 * max(1.0, abs(z))


# -------------------------------------------------------------- radiation reaction
def test_fluxes_reach_the_peters_limit():
    # dp/dt -> -64/5 (1-e^2)^{3/2} (1 + 7/8 e^2) / p^3, de/dt -> -304/15 e (...)(1 + 121/304 e^2) / p^4,
    # with 1PN corrections O(1/p): p (ratio - 1) stays bounded and settles
    prev = None
    for p in (100.0, 200.0, 400.0):
        e = 0.3
        d = np.zeros(3)
        O.o_param_derivs(p, e, 0.8, 0.5, ol._ptr(d))
        rp = d[0] / (-(64 / 5) * (1 - e * e) ** 1.5 * (1 + 7 / 8 * e * e) / p ** 3)
        re = d[1] / (-(304 / 15) * e * (1 - e * e) ** 1.5 * (1 + 121 / 304 * e * e) / p ** 4)
        assert abs(rp - 1) * p < 3 and abs(re - 1) * p < 4
        if prev is not None:
            assert abs(rp - 1) < abs(prev - 1)
        prev = rp


# ------------------------------------------------------------------ AK PN maps
def test_ak_rates_newtonian_limit():
    for e in (0.0, 0.3, 0.6):
        for v in (1e-3, 1e-2):
            lead = 32 / 5 * (1 - e * e) ** 1.5 * v ** 9 * (1 + 7 / 8 * e * e)
            r = O.o_dvdt(v, e, 0.7, 2.0, 3.0, 0.4) / (lead * 2.0 / 9.0)
            assert abs(r - 1) < 3 * v
            if e > 0:
                lead = -(304 / 15) * v ** 8 * e * (1 - e * e) ** 1.5 * (1 + 121 / 304 * e * e)
                r = O.o_dedt(v, e, 0.7, 2.0, 3.0, 0.4) / (lead * 2.0 / 9.0)
                assert abs(r - 1) < 3 * v


def test_no_nodal_precession_without_spin():
    rng = np.random.default_rng(24)
    for v, e, Y in zip(rng.uniform(0.05, 0.45, 50), rng.uniform(0, 0.8, 50), rng.uniform(-1, 1, 50)):
        assert O.o_dthetadm(v, e, Y, 0.0) == pytest.approx(O.o_dphidm(v, e, Y, 0.0), rel=1e-15)


def test_periastron_and_lense_thirring_precession():
    for e in (0.0, 0.4):
        for Y in (0.3, 0.9):
            v = 1e-3
            # periastron advance: Omega_phi/Omega_r - 1 -> 3 v^2
            k = O.o_dphidm(v, e, Y, 0.0) / O.o_drdm(v, e, Y, 0.0) - 1
            assert k / (3 * v * v) == pytest.approx(1, abs=1e-4)
            # Lense-Thirring nodal precession: (Omega_phi - Omega_theta)/Omega_r -> 2 q v^3
            q = 0.7
            lt = (O.o_dphidm(v, e, Y, q) - O.o_dthetadm(v, e, Y, q)) / O.o_drdm(v, e, Y, q)
            assert lt / (2 * q * v ** 3) == pytest.approx(1, abs=1e-3)


def test_omega_phi_kepler_limit_and_retrograde_symmetry():
    for v in (1e-3, 1e-2):
        assert O.o_omegaphi(v, 0.0, 0.8, 0.0, 2.0) * 2.0 / v ** 3 == pytest.approx(1, abs=3 * v * v)
    # retrograde orbits use (-cos iota, -s)
    assert O.o_omegaphi(0.3, 0.2, -0.6, 0.5, 1.0) == O.o_omegaphi(0.3, 0.2, 0.6, -0.5, 1.0)


def test_mean_anomaly_from_true_anomaly():
    # Phi0 = E - e sin E with tan(E/2) = sqrt((1-e)/(1+e)) tan(psi/2)
    rng = np.random.default_rng(25)
    for e, psi in zip(rng.uniform(0, 0.9, 100), rng.uniform(-np.pi, np.pi, 100)):
        E = 2 * np.arctan(np.sqrt((1 - e) / (1 + e)) * np.tan(psi / 2))
        got = O.o_parang_phi0(e, psi)
        d = (got - (E - e * np.sin(E)) + np.pi) % (2 * np.pi) - np.pi
        assert abs(d) < 1e-12


# ------------------------------------------------------------------ mode sum
def _kepler_quadrupole(e, Phi):
    """Second mean-anomaly derivatives of the Keplerian quadrupole (a = 1):
    (X^2-Y^2, 2XY, X^2+Y^2) with X = cos E - e, Y = sqrt(1-e^2) sin E."""
    E = brentq(lambda E: E - e * np.sin(E) - Phi, Phi - 2, Phi + 2, xtol=1e-15, rtol=1e-15)
    D, Dp = 1 - e * np.cos(E), e * np.sin(E)
    s = np.sqrt(1 - e * e)
    X, Xp, Xpp = np.cos(E) - e, -np.sin(E), -np.cos(E)
    Y, Yp, Ypp = s * np.sin(E), s * np.cos(E), -s * np.sin(E)

    def d2(fp, fpp):
        return (fpp * D - fp * Dp) / D ** 3

    return np.array([d2(2 * (X * Xp - Y * Yp), 2 * (Xp * Xp + X * Xpp - Yp * Yp - Y * Ypp)),
                     d2(2 * (Xp * Y + X * Yp), 2 * (Xpp * Y + 2 * Xp * Yp + X * Ypp)),
                     d2(2 * (X * Xp + Y * Yp), 2 * (Xp * Xp + X * Xpp + Yp * Yp + Y * Ypp))]), 1 / D - 1


@pytest.mark.parametrize("e", [0.05, 0.3, 0.6])
def test_mode_coefficients_sum_to_the_keplerian_quadrupole(e):
    # Peters-Mathews: sum_n (a_n, b_n, c_n) = 1/2 d^2/dPhi^2 (X^2-Y^2, 2XY, X^2+Y^2) and
    # sum_n c_n = a/r - 1 (Kapteyn); the series is truncated where J_n(ne) < 1e-30
    abc = np.zeros(3)
    for Phi in (0.3, 1.7, 2.9, 4.0, 5.9):
        S = np.zeros(3)
        for n in range(1, 121):
            O.o_mode_abc(n, e, 1.0, n * Phi, ol._ptr(abc))
            S += abc
        quad_, c_closed = _kepler_quadrupole(e, Phi)
        np.testing.assert_allclose(S, 0.5 * quad_, rtol=1e-11, atol=1e-12)
        assert S[2] == pytest.approx(c_closed, rel=1e-11, abs=1e-13)


def test_mode_amplitude_is_linear():
    a1, a2 = np.zeros(3), np.zeros(3)
    O.o_mode_abc(7, 0.4, 1.0, 2.2, ol._ptr(a1))
    O.o_mode_abc(7, 0.4, 3.5, 2.2, ol._ptr(a2))
    np.testing.assert_allclose(a2, 3.5 * a1, rtol=1e-15)


# --------------------------------------------------------- frame rotation
def test_rotation_coefficients_are_a_rotation_by_twice_the_polarisation_angle():
    rng = np.random.default_rng(26)
    rot = np.zeros(4)
    for _ in range(100):
        iota, thS, phS, thK, phK, alp = rng.uniform(0.1, 3.0), rng.uniform(0.1, 3.0), rng.uniform(0, 6.28), \
            rng.uniform(0.1, 3.0), rng.uniform(0, 6.28), rng.uniform(0, 6.28)
        O.o_rotcoeff(ol._ptr(rot), iota, thS, phS, thK, phK, alp)
        n = np.array([np.sin(thS) * np.cos(phS), np.sin(thS) * np.sin(phS), np.cos(thS)])
        S = np.array([np.sin(thK) * np.cos(phK), np.sin(thK) * np.sin(phK), np.cos(thK)])
        # L at angle iota from S, precessing with phase alp about S
        ex = np.array([np.cos(thK) * np.cos(phK), np.cos(thK) * np.sin(phK), -np.sin(thK)])
        ey = np.cross(S, ex)
        L = np.cos(iota) * S + np.sin(iota) * (-np.cos(alp) * ex - np.sin(alp) * ey)
        u, w = np.cross(n, L), np.cross(n, S)
        ang = np.arctan2(np.dot(n, np.cross(w, u)), np.dot(u, w))
        np.testing.assert_allclose(rot, [np.cos(2 * ang), np.sin(2 * ang), -np.sin(2 * ang), np.cos(2 * ang)],
                                   atol=1e-12)


# ------------------------------------------------------ spline, fit, map
def test_natural_spline_against_scipy():
    rng = np.random.default_rng(27)
    x = np.concatenate([[0.0], np.cumsum(rng.uniform(0.1, 1.0, 60))])
    y = np.sin(x) + rng.normal(scale=0.1, size=x.size)
    c = np.zeros(x.size)
    assert O.o_spline_natural(ol._ptr(x), ol._ptr(y), x.size, ol._ptr(c)) == 0
    cs = CubicSpline(x, y, bc_type="natural")
    xs = np.linspace(x[0], x[-1], 997)
    got = np.array([O.o_spline_eval(ol._ptr(x), ol._ptr(y), ol._ptr(c), x.size, t) for t in xs])
    np.testing.assert_allclose(got, cs(xs), rtol=1e-13, atol=1e-13)


def test_anchored_quadratic_fit_against_lstsq():
    rng = np.random.default_rng(28)
    x = np.sort(rng.uniform(0, 5, 40))
    y = 1.5 + 0.3 * (x - x[0]) - 0.07 * (x - x[0]) ** 2 + rng.normal(scale=1e-3, size=x.size)
    coeff = np.zeros(2)
    assert O.o_polyfit(ol._ptr(coeff), ol._ptr(x), ol._ptr(y), x.size) == 0
    dx = x[1:] - x[0]
    ref = np.linalg.lstsq(np.column_stack([dx, dx * dx]), y[1:] - y[0], rcond=None)[0]
    np.testing.assert_allclose(coeff, ref, rtol=1e-11)


def test_frequency_map_solves_the_system_like_minpack_hybr():
    # ParMap: Omega_Kerr(p, e, iota) M_map = Omega_AK(v, e, iota, s_map) M; the oracle's Powell hybrid
    # must land on the same root as scipy's MINPACK hybr, with the reference's residual stop (1e-6)
    for p, e, iota, s, M in ((8.0, 0.2, 0.5, 0.5, 1e6), (12.0, 0.45, 1.0, 0.9, 3e5), (7.5, 0.05, 2.4, 0.3, 5e6)):
        g = ol.o_iekg(p, e, np.cos(iota), s)
        Om = np.ascontiguousarray(g[13:16])
        m = np.zeros(3)
        it = O.o_parmap(ol._ptr(m), ol._ptr(Om), p, M, s, e, iota)
        assert it >= 0
        f = np.zeros(3)
        O.o_sol_fun(ol._ptr(m), ol._ptr(Om), M, e, iota, ol._ptr(f))
        assert np.sum(np.abs(f)) < 1e-6

        def fun(x):
            xx = np.ascontiguousarray(x)
            out = np.zeros(3)
            O.o_sol_fun(ol._ptr(xx), ol._ptr(Om), M, e, iota, ol._ptr(out))
            return out

        sol = root(fun, [1 / np.sqrt(p), M, s], method="hybr")
        assert sol.success
        np.testing.assert_allclose(m, sol.x, rtol=2e-6)
