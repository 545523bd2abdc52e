"""Pins of the oracle's AK radiation-reaction rates and of its PNevolution
glue (o_dvdt/o_dedt: KSParMap.cc:86-144; o_pn_knots: AAK.cc:27-209;
o_pn_dense i_plunge/i_buffer: AAK.cc:239-243).

* dv/dt at e = 0 is derived here, independently of the oracle's polynomial,
  from the test-mass energy flux of a circular equatorial Kerr orbit through
  3PN (Tagoshi, Shibata, Tanaka & Sasaki 1996, including the 1.5PN and 2.5PN
  spin-orbit, the 2PN spin-spin and the 3PN logarithmic terms) and the exact
  Kerr circular-orbit energy E(r): with v = r^(-1/2) (the map variable) and
  x = (M Omega)^(2/3), dv/dt = (16/5)(mu/M^2) v^3 x^5 F(x) / (dE/dr),
  expanded in v to the order of the oracle's series.
* the AK evolution converges at second order (Adams-Bashforth 2) towards an
  accurate ODE solution of the same rates, and the anchored quadratic fits
  reproduce quadratic mass and spin drifts exactly.
* the plunge index follows the reference's detection rule, checked knot by
  knot with the oracle's Kerr geodesic (itself pinned against the compiled
  reference) instead of the oracle's strided scan and bisection.
* i_plunge, i_buffer (ten azimuthal periods) and the backward knot times
  follow their closed forms.
"""
import os
import sys

import numpy as np
import pytest
import sympy as sp
from scipy.integrate import solve_ivp

import oracle_lib as ol

O = ol.oracle()
MSUN_SEC = 6.67259e-11 * 1.98892e30 / (299792458.0 * 299792458.0 * 299792458.0)  # Constants.h


# --------------------------------------------------------------- dv/dt, e = 0
def _series_mul(a, b, n):
    c = [sp.Integer(0)] * n
    for i in range(n):
        for j in range(n - i):
            c[i + j] += a[i] * b[j]
    return [sp.expand(x) for x in c]


def _series_pow(a, alpha, n):
    """(1 + a[1] v + a[2] v^2 + ...)^alpha, truncated at v^(n-1)."""
    rest = [sp.Integer(0)] + list(a[1:])
    res = [sp.Integer(1)] + [sp.Integer(0)] * (n - 1)
    term = list(res)
    coef = sp.Integer(1)
    for k in range(1, n):
        coef = coef * (alpha - k + 1) / k
        term = _series_mul(term, rest, n)
        res = [sp.expand(res[i] + coef * term[i]) for i in range(n)]
    return res


def _poly(d, n):
    a = [sp.Integer(0)] * n
    for k, v in d.items():
        a[k] = v
    return a


def circular_dvdt_series():
    """Coefficients c_0..c_6 of (dv/dt) / ((32/5)(mu/M^2) v^9) for a circular
    equatorial orbit (prograde for q > 0, retrograde as q < 0), in powers of
    v = r^(-1/2); c_6 holds log(v) through the symbol Lv."""
    n = 10
    q, Lv = sp.symbols("q Lv")
    R = sp.Rational
    # E(r) = (1 - 2/r + q r^-3/2) / sqrt(1 - 3/r + 2 q r^-3/2), r = v^-2
    E = _series_mul(_poly({0: 1, 2: -2, 3: q}, n), _series_pow(_poly({0: 1, 2: -3, 3: 2 * q}, n), R(-1, 2), n), n)
    dE_over_v = [sp.expand((i + 2) * E[i + 2]) for i in range(n - 2)] + [0, 0]  # (dE/dv) / v
    # dE/dr = -(v^3/2) dE/dv, so the ratio is -(1 + q v^3)^(-10/3) F v / (dE/dv)
    inv = _series_pow([sp.expand(-c) for c in dE_over_v], -1, n)
    g, pi = sp.EulerGamma, sp.pi
    flux = {0: 1, 2: -R(1247, 336), 3: 4 * pi - R(11, 4) * q, 4: -R(44711, 9072) + R(33, 16) * q ** 2,
            5: -R(8191, 672) * pi - R(59, 16) * q,
            6: R(6643739519, 69854400) - R(1712, 105) * g + R(16, 3) * pi ** 2 - R(3424, 105) * sp.log(2)
            - R(1712, 105) * Lv - R(65, 6) * pi * q + R(611, 504) * q ** 2}
    # F in powers of sqrt(x) = v (1 + q v^3)^(-1/3); log x = 2 log v + O(v^3) only enters beyond v^6
    F = [sp.Integer(0)] * n
    for k, ck in flux.items():
        t = _series_pow(_poly({0: 1, 3: q}, n), R(-k, 3), n)
        for i in range(n - k):
            F[i + k] += ck * t[i]
    F = [sp.expand(x) for x in F]
    ratio = _series_mul(_series_mul(_series_pow(_poly({0: 1, 3: q}, n), R(-10, 3), n), F, n), inv, n)
    return ratio[:7], q, Lv


def test_circular_dvdt_matches_the_test_mass_flux_through_3pn():
    coef, q, Lv = circular_dvdt_series()
    assert coef[0] == 1 and coef[1] == 0
    assert coef[2] == sp.Rational(-743, 336)
    for qq in (0.0, 0.3, 0.6, 0.95, -0.4, -0.9):
        c = [sp.lambdify((q, Lv), ck) for ck in coef]
        for Y in (1.0, -1.0):
            # a retrograde orbit is the prograde one about spin -q
            qe = qq * Y
            for v in (0.02, 0.05, 0.1, 0.2, 0.3):
                lead = 32 / 5 * v ** 9
                got = O.o_dvdt(v, 0.0, Y, 1.0, 1.0, qq) / lead
                exp = sum(c[k](qe, np.log(v)) * v ** k for k in range(7))
                # a coefficient wrong by d at order k moves the value by d v^k
                assert abs(got - exp) < 1e-14, (qq, Y, v, got, exp)


def test_circular_dvdt_coefficients_one_by_one():
    # the same comparison order by order: the oracle's series extracted by a
    # least-squares fit in (v^2, v^3, v^4, v^5, v^6, v^6 log v) at fixed spin
    coef, q, Lv = circular_dvdt_series()
    v = np.linspace(0.03, 0.45, 60)
    A = np.column_stack([v ** 2, v ** 3, v ** 4, v ** 5, v ** 6, v ** 6 * np.log(v)])
    for qq in (0.0, 0.7, -0.7):
        y = np.array([O.o_dvdt(x, 0.0, 1.0, 1.0, 1.0, qq) / (32 / 5 * x ** 9) - 1 for x in v])
        fit = np.linalg.lstsq(A, y, rcond=None)[0]
        want = [float(coef[k].subs({q: qq, Lv: 0})) for k in (2, 3, 4, 5, 6)]
        want.append(float(sp.diff(coef[6], Lv)))
        np.testing.assert_allclose(fit, want, rtol=1e-6, atol=1e-6)


# ---------------------------------------------------------- AK evolution + fits
def _synthetic_sparse(steps, T, mu, M, s, e0, v0, coslam):
    """Sparse inputs whose (e, v) are an accurate solution of the AK rates and
    whose mass and spin drift quadratically in time."""
    par = np.array([mu * MSUN_SEC, M * MSUN_SEC, s, e0, np.arccos(coslam), 0, 0, 0, 0, 0, 0, 0])

    def rhs(t, y):
        v, e = y
        return [O.o_dvdt(v, e, coslam, par[0], par[1], s), O.o_dedt(v, e, coslam, par[0], par[1], s)]

    sol = solve_ivp(rhs, (0, 2.2 * T), [v0, e0], method="DOP853", rtol=1e-13, atol=1e-15, dense_output=True)
    dt_map = np.linspace(0, T, steps)
    v_map, e_traj = sol.sol(dt_map)
    M_map = M * (1 + 3e-6 * dt_map / T - 2e-6 * (dt_map / T) ** 2)
    s_map = s + 1e-4 * dt_map / T + 3e-5 * (dt_map / T) ** 2
    return par, e_traj, v_map, M_map, s_map, dt_map, sol


def _knots_from_sparse(par, e_traj, v_map, M, M_map, s, s_map, dt_map, length, dt):
    steps = dt_map.size
    j0 = O.o_pn_j0(dt, length, ol._ptr(dt_map), steps)
    cap = 2 * j0 + 8
    kn = np.zeros(5 * cap)
    info = np.zeros(9, dtype=np.int32)
    K = O.o_pn_knots(dt, length, ol._ptr(par), ol._ptr(e_traj), ol._ptr(v_map), M, ol._ptr(M_map), s,
                     ol._ptr(s_map), ol._ptr(dt_map), steps, 0, *[ol._ptr(kn[k * cap:]) for k in range(5)], cap,
                     info.ctypes.data_as(ol._pi))
    assert K > 0
    return kn.reshape(5, cap)[:, :K], info


def test_ak_evolution_converges_at_second_order_and_fits_are_exact():
    M, mu, s, coslam = 1e5, 10.0, 0.6, np.cos(0.5)
    T = 3.0e5  # seconds: v climbs from 0.300 to ~0.309 over the sparse span
    length, dt = 40000, 15.0  # 6e5 s of samples: knots run to twice the sparse span
    errs = []
    for steps in (25, 50, 100):
        par, e_traj, v_map, M_map, s_map, dt_map, sol = _synthetic_sparse(steps, T, mu, M, s, 0.3, 0.30, coslam)
        kn, info = _knots_from_sparse(par, np.ascontiguousarray(e_traj), np.ascontiguousarray(v_map), M, M_map, s,
                                      s_map, dt_map, length, dt)
        assert info[7] == 0  # no plunge, no clamp
        t, e, v, Mk, Sk = kn
        # knot times: the sparse times, then uniform steps of dt_large
        np.testing.assert_array_equal(t[:steps], dt_map)
        np.testing.assert_allclose(np.diff(t[steps - 1:]), dt_map[1], rtol=1e-12)
        ve, ee = sol.sol(t)
        errs.append(max(np.max(np.abs(v - ve)), np.max(np.abs(e - ee))))
        # quadratic drifts through the anchored fits, converted to seconds
        Mq = M * (1 + 3e-6 * t / T - 2e-6 * (t / T) ** 2) * MSUN_SEC
        Sq = s + 1e-4 * t / T + 3e-5 * (t / T) ** 2
        np.testing.assert_allclose(Mk, Mq, rtol=1e-13)
        np.testing.assert_allclose(Sk, Sq, rtol=1e-11)
    # halving the step divides the error by ~4 (first order would give ~2)
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.3 < r1 < 5.0 and 3.3 < r2 < 5.0, errs


# ---------------------------------------------------------- plunge and buffer
def _isco(a):
    """Prograde and retrograde circular equatorial ISCO radii (Bardeen, Press
    & Teukolsky 1972)."""
    z1 = 1 + (1 - a * a) ** (1 / 3) * ((1 + a) ** (1 / 3) + (1 - a) ** (1 / 3))
    z2 = np.sqrt(3 * a * a + z1 * z1)
    d = np.sqrt((3 - z1) * (3 + z1 + 2 * z2))
    return 3 + z2 - d, 3 + z2 + d


def _p_of_knot(v, e, S, M, coslam, Mphys):
    return (1 - e * e) * (O.o_omegaphi(v, e, coslam, S, M) * Mphys * MSUN_SEC) ** (-2 / 3)


def _unstable(p, e, coslam, a):
    try:
        g = ol.o_iekg(p, e, coslam, a, freq=False)
    except RuntimeError:
        return True
    return g[3] == -1 or g[0] > 1


def _production_sources(count, max_span):
    """Production-prior draws (1-yr-class, dt = 15 s, N = 2^21) that plunge
    within max_span knots of the start."""
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    import bench
    rng = np.random.default_rng(11)
    out = []
    while len(out) < count:
        pars = bench.draw_params(rng)
        try:
            r = ol.o_knots(pars, 1.0, 2 ** 21, 15.0)
        except RuntimeError:
            continue
        i = r["info"]
        if i[7] and i[3] - i[1] < max_span:
            out.append((pars, r))
    return out


def test_plunge_index_by_brute_force():
    # The reference's rule (AAK.cc:149-173), checked knot by knot: the plunge
    # is detected at the first knot j whose p lies inside the prograde ISCO,
    # or, on knots every radiation-reaction time tau = M^2/mu after the start,
    # whose p lies inside the retrograde ISCO with an unstable geodesic; the
    # plunge knot is then the first unstable knot of the tau before it (the
    # window's first knot is taken as stable without a test).  Without a
    # detection the same refinement runs from the last check point to the end
    # of the knots.  Eccentric separatrices lie outside both ISCOs, so the
    # plunge knot can come after the first unstable knot of the whole scan.
    for pars, r in _production_sources(20, 30000):
        i0, j0, j_min, j_max, j_max_temp, j_start = r["info"][:6]
        coslam, a, Mphys = np.cos(pars[0]), r["S_phys"], r["M_phys"]
        lso_min, lso_max = _isco(a)
        # one radiation-reaction time in knots; tau / dt_large = 1 / dt_fit is
        # 99 up to rounding, so the C expression's own order is kept
        j_rr = int(MSUN_SEC * Mphys * MSUN_SEC * Mphys / r["par"][0] / r["dt_map"][1]) + 1
        p = np.array([_p_of_knot(r["v"][j - j_start], r["e"][j - j_start], r["S"][j - j_start],
                                 r["M"][j - j_start], coslam, Mphys) for j in range(j0, j_max_temp + 1)])
        unstable = {}

        def bad(j):
            if j not in unstable:
                unstable[j] = _unstable(p[j - j0], r["e"][j - j_start], coslam, a)
            return unstable[j]

        detect = None
        for j in range(j0, j_max_temp + 1):
            if p[j - j0] < lso_min or (j > j0 and (j - j0) % j_rr == 0 and p[j - j0] < lso_max and bad(j)):
                detect = j
                break
        if detect is not None:
            lo, hi = max(detect - j_rr, j0), min(detect + 1, j_max_temp)
        else:
            # no detection: the refinement still runs from the last check
            # point inside the retrograde ISCO up to the end of the knots
            checks = [j for j in range(j0 + j_rr, j_max_temp + 1, j_rr) if p[j - j0] < lso_max]
            lo, hi = (max(checks[-1] - j_rr, j0) if checks else j0), j_max_temp
        # bisection from (lo, hi] with lo taken as stable and hi as unstable
        first = next((j for j in range(lo + 1, hi + 1) if bad(j)), hi)
        assert (j_min, j_max) == (first - 1, first), (pars, j_min, j_max, first)


def test_plunge_sample_and_buffer_closed_forms():
    # i_plunge is the sample at or before the last stable knot's time; the
    # buffer spans ten azimuthal periods 2 pi / Omega_phi there (+1 sample)
    length, dt = 2 ** 16, 30.0
    rng = np.random.default_rng(12)
    seen = 0
    while seen < 6:
        pars = [0.6, rng.uniform(0.1, 0.9), rng.uniform(7.5, 9.0), rng.uniform(0.05, 0.4), 10 ** rng.uniform(5, 5.5),
                rng.uniform(20, 50), 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
        try:
            r = ol.o_knots(pars, 1.0, length, dt)
            w = ol.o_aak_wave(pars, 1.0, length, dt, want_dense=True)
        except RuntimeError:
            continue
        if not r["info"][7]:
            continue
        seen += 1
        i0, j_min, j_start = r["info"][0], r["info"][2], r["info"][5]
        t_last_stable = r["t"][j_min - j_start]
        ip, ib = w["i_plunge"], w["i_buffer"]
        t = w["dense_t"]
        assert t[ip] <= t_last_stable < t[ip + 1]
        assert (ip - i0) * dt <= t_last_stable < (ip - i0 + 1) * dt
        Om = O.o_omegaphi(w["dense_v"][ip], w["dense_e"][ip], np.cos(pars[0]), w["dense_S"][ip], w["dense_M"][ip])
        periods = 10 * 2 * np.pi / Om / dt
        assert ib - 1 <= periods < ib


def test_backward_knots_step_back_from_the_start():
    pars = [0.5, 0.5, 8.0, 0.1, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    length, dt = 2 ** 18, 120.0  # 1 yr: the source plunges after ~322 d
    r = ol.o_knots(pars, 1.0, length, dt, backint=True)
    i0, j0, j_min, j_max, j_max_temp, j_start, j_end, plunged = r["info"][:8]
    assert plunged and j_start < j0 and j_end == j_max
    dt_large = r["dt_map"][1]
    k = np.arange(j_start, j0)
    np.testing.assert_array_equal(r["t"][:j0 - j_start], dt_large * (k - j0))
    # the forward knots keep their forward times
    np.testing.assert_array_equal(r["t"][j0 - j_start:j0 - j_start + r["steps"]], r["dt_map"])
    # the waveform then ends at the plunge: t_end = min(t_fit[j_max], T)
    assert r["t"][-1] <= length * dt
