// aak_geodesic.hpp -- bound Kerr geodesics and kludge radiation reaction.
//
//   Geodesic<CX>  constants (E, Lz, Q) of an inclined eccentric orbit by a
//                 2-D Newton iteration, radial roots by Laguerre's method, and
//                 the three fundamental frequencies (IEKG.cc:19-226,
//                 IEKGMNewt.cc:7-48; frequencies after Schmidt 2002)
//   circular_constants  circular orbit at fixed cos(iota) (CKG.cc:9-74)
//   kludge_rates        Gair-Glampedakis fluxes -> (pdot, edot, d cos(iota)/dt)
//                       (GKR.cc:28-97,195-417)
//
// CX is the complex type used by the root finder: std::complex<double> on the
// host (so that division/abs/sqrt go through the same runtime routines as the
// reference build), aak::DevComplex on the device.
#pragma once

#include "aak_math.hpp"

namespace aak {

// Complex arithmetic for device code (Smith's division with rescaling is not
// needed here: the quartic's coefficients are O(1..1e4)).  Kept in its own
// namespace so that its abs/sqrt overloads are found only by argument-dependent
// lookup and never shadow the real ones.
namespace dc {
struct DevComplex {
  double re, im;
  AAK_HD DevComplex() : re(0), im(0) {}
  AAK_HD DevComplex(double r, double i = 0.0) : re(r), im(i) {}
  AAK_HD double real() const { return re; }
  AAK_HD double imag() const { return im; }
};
AAK_HD DevComplex operator+(DevComplex a, DevComplex b) { return {a.re + b.re, a.im + b.im}; }
AAK_HD DevComplex operator-(DevComplex a, DevComplex b) { return {a.re - b.re, a.im - b.im}; }
AAK_HD DevComplex operator*(DevComplex a, DevComplex b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
AAK_HD DevComplex operator*(double s, DevComplex a) { return {s * a.re, s * a.im}; }
AAK_HD DevComplex operator/(DevComplex a, DevComplex b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return {(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.im + b.re * r;
  return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}
AAK_HD double abs(DevComplex a) { return hypot(a.re, a.im); }
AAK_HD DevComplex sqrt(DevComplex a) {
  if (a.re == 0.0 && a.im == 0.0) return {0.0, a.im};
  const double m = hypot(a.re, a.im);
  double t = ::sqrt(0.5 * (m + fabs(a.re)));
  if (a.re >= 0.0) return {t, a.im / (2.0 * t)};
  return {fabs(a.im) / (2.0 * t), (a.im >= 0.0) ? t : -t};
}
}  // namespace dc
using dc::DevComplex;

// Complex quotient by Smith's algorithm in the operation order of the GNU
// runtime's __divdc3 (the routine the reference build calls for complex
// division); the LLVM runtime's scaled Annex-G division rounds differently.
template <class CX>
AAK_HD CX cdiv(const CX &num, const CX &den) {
  const double a = num.real(), b = num.imag(), c = den.real(), d = den.imag();
  if (fabs(c) < fabs(d)) {
    const double r = c / d, q = (c * r) + d;
    return CX(((a * r) + b) / q, ((b * r) - a) / q);
  }
  const double r = d / c, q = (d * r) + c;
  return CX(((b * r) + a) / q, (b - (a * r)) / q);
}

// Laguerre's method with deflation and polishing for a real quartic (the
// method of Press et al.; the reference uses NRzroot.cc/NRLaguer.cc).
template <class CX>
struct QuarticRoots {
  AAK_HD static int laguer(const CX *a, int m, CX *x) {
    const double frac[16] = {0.0, 0.5, 0.25, 0.75, 0.13, 0.38, 0.62, 0.88, 1.0, 0, 0, 0, 0, 0, 0, 0};
    for (int iter = 1; iter <= 300; iter++) {
      CX b = a[m], d = CX(0.0, 0.0), f = CX(0.0, 0.0);
      double err = abs(b);
      const double abx = abs(*x);
      for (int j = m - 1; j >= 0; j--) {
        f = (*x) * f + d;
        d = (*x) * d + b;
        b = (*x) * b + a[j];
        err = abs(b) + abx * err;
      }
      err *= 1.0e-15;
      if (abs(b) <= err) return 0;
      const CX g = cdiv(d, b);
      const CX g2 = g * g;
      const CX h = g2 - cdiv(CX(2.0 * f.real(), 2.0 * f.imag()), b);
      const CX sq = sqrt(((double)(m - 1)) * (((double)m) * h - g2));
      CX gp = g + sq;
      const CX gm = g - sq;
      const double abp = abs(gp), abm = abs(gm);
      if (abp < abm) gp = gm;
      CX dx;
      if (dmax(abp, abm) > 0.0) {
        dx = cdiv(CX((double)(float)m, 0.0), gp);
      } else {
        dx = exp(log(1 + abx)) * CX((double)cosf((float)iter), (double)sinf((float)iter));
      }
      const CX x1 = (*x) - dx;
      if (x->real() == x1.real() && x->imag() == x1.imag()) return 0;
      if (iter % 20) *x = x1;
      else *x = *x - frac[iter / 20] * dx;
    }
    return -1;
  }
  // Real parts of the four roots in descending order.
  AAK_HD static int solve(const double *coef, double *r_desc) {
    CX a[5], ad[5], roots[5];
    for (int i = 0; i < 5; i++) a[i] = ad[i] = CX(coef[i], 0.0);
    for (int j = 4; j >= 1; j--) {
      CX x = CX(0.0, 0.0);
      if (laguer(ad, j, &x)) return -1;
      if (fabs(x.imag()) <= 2.0 * 1.e-15 * fabs(x.real())) x = CX(x.real(), 0.0);
      roots[j] = x;
      CX b = ad[j];
      for (int jj = j - 1; jj >= 0; jj--) {
        const CX c = ad[jj];
        ad[jj] = b;
        b = x * b + c;
      }
    }
    for (int j = 1; j <= 4; j++)
      if (laguer(a, 4, &roots[j])) return -1;
    for (int j = 2; j <= 4; j++) {
      const CX x = roots[j];
      int i;
      for (i = j - 1; i >= 1; i--) {
        if (roots[i].real() <= x.real()) break;
        roots[i + 1] = roots[i];
      }
      roots[i + 1] = x;
    }
    for (int i = 0; i < 4; i++) r_desc[i] = roots[4 - i].real();
    return 0;
  }
};

// Sigma^2 (dr/dtau)^2 of a Kerr geodesic (GKG.cc:8-16).
AAK_HD double radial_potential(double r, double a, double E, double Lz, double Q) {
  const double t1 = E * (r * r + a * a) - a * Lz;
  const double lma = Lz - a * E;
  const double t2 = Kerr::delta(r, a) * (r * r + lma * lma + Q);
  return t1 * t1 - t2;
}

template <class CX>
struct Geodesic {
  double p, ecc, cosiota, a;
  double peri, ap, tansq, tani;
  double E, Lz, Q, zplus, zminus, bzplus, beta_w;
  double r[4];
  int stable;  // 1 stable, 0 marginal, -1 unstable, 2 unclassified

  // Newton on (E, Lz) with a 2x2 Crout factorisation using implicit row
  // scaling (the scheme of the reference's mnewt/ludcmp/lubksb), at most 200
  // iterations.  An iteration depends on the state x alone, so once x equals
  // bit for bit the state of P iterations earlier (near-circular orbits often
  // never meet the tolerance and settle into such a cycle of period <= ~40),
  // every remaining iteration repeats that cycle: the state the 200th
  // iteration would leave is read from the recorded cycle.
  AAK_HD int newton(double *x) {
    constexpr int kRing = 64;
    double ring[kRing][2];  // state at the start of iterations k-63..k
    for (int k = 1; k <= 200; k++) {
      for (int P = 1; P <= kRing - 1 && P < k; P++) {
        const double *o = ring[(k - P) % kRing];
        if (bits_equal(o[0], x[0]) && bits_equal(o[1], x[1])) {
          const int j = k - P + (201 - k) % P;  // state at the start of "iteration 201"
          x[0] = ring[j % kRing][0];
          x[1] = ring[j % kRing][1];
          return 0;
        }
      }
      ring[k % kRing][0] = x[0];
      ring[k % kRing][1] = x[1];
      const double e = x[0], lz = x[1];
      const double ti2 = tani * tani, ci2 = cosiota * cosiota, q = lz * lz * ti2;
      const double f0 = radial_potential(peri, a, e, lz, q);
      const double f1 = radial_potential(ap, a, e, lz, q);
      const double aelz = a * e - lz;
      double j00 = 2. * peri * (2. * a * aelz + peri * (a * a * e + peri * peri * e));
      double j01 = -2. * a * a * lz * ti2 + 2. * peri * (2. * (lz * ti2 - aelz) - peri * lz / ci2);
      double j10 = 2. * ap * (2. * a * aelz + ap * (a * a * e + ap * ap * e));
      double j11 = -2. * a * a * lz * ti2 + 2. * ap * (2. * (lz * ti2 - aelz) - ap * lz / ci2);
      double errf = 0.0;
      errf += fabs(f0);
      errf += fabs(f1);
      if (errf <= 1.e-14) return 0;
      double b0 = -f0, b1 = -f1;
      double s0 = 0.0, s1 = 0.0;
      if (fabs(j00) > s0) s0 = fabs(j00);
      if (fabs(j01) > s0) s0 = fabs(j01);
      if (fabs(j10) > s1) s1 = fabs(j10);
      if (fabs(j11) > s1) s1 = fabs(j11);
      if (s0 == 0.0 || s1 == 0.0) return -1;
      const double vv0 = 1.0 / s0, vv1 = 1.0 / s1;
      // pivot choice for column 0 (">=" keeps the later row on ties)
      const double d0 = vv0 * fabs(j00), d1 = vv1 * fabs(j10);
      bool swap = (d1 >= d0);
      if (swap) {
        double t = j00; j00 = j10; j10 = t;
        t = j01; j01 = j11; j11 = t;
        t = b0; b0 = b1; b1 = t;
      }
      if (j00 == 0.0) j00 = 1.0e-20;
      j10 *= 1.0 / j00;
      j11 = j11 - j10 * j01;
      if (j11 == 0.0) j11 = 1.0e-20;
      if (b0 != 0.0) b1 -= j10 * b0;
      b1 = b1 / j11;
      double s = b0;
      s -= j01 * b1;
      b0 = s / j00;
      double errx = 0.0;
      errx += fabs(b0);
      x[0] += b0;
      errx += fabs(b1);
      x[1] += b1;
      if (errx <= 1.e-14) return 0;
    }
    return 0;
  }

  // roots: 0 = constants only, 1 = also radial roots and stability flag
  AAK_HD int init(double p_, double e_, double ci_, double a_, int roots) {
    p = p_; ecc = e_; cosiota = ci_; a = a_;
    if (ecc == 0) return -1;
    peri = p / (1. + ecc);
    ap = p / (1. - ecc);
    tansq = (1. - cosiota) * (1. + cosiota) / (cosiota * cosiota);
    tani = sqrt(tansq);
    const double semimaj = peri / (1. - ecc);
    double epro = Kerr::e_pro(semimaj, a), eret = Kerr::e_ret(semimaj, a);
    if (is_nan(epro) || epro > 1. - 1.e-6) epro = 1. - 1.e-6;
    if (is_nan(eret) || eret > 1.) eret = 1.;
    const double eg = 0.5 * (cosiota + 1.) * epro + 0.5 * (1. - cosiota) * eret;
    double x[2] = {eg, cosiota * sqrt((1. - ecc) * (1. + ecc) / (2. - 2. * eg))};
    if (newton(x)) return -1;
    E = x[0];
    Lz = x[1];
    Q = Lz * Lz * tansq;
    const double alpha_w = Q + Lz * Lz;
    beta_w = a * a * (1. - E * E);
    const double tmp = alpha_w + beta_w;
    if (beta_w > 1.e-15) {
      zplus = (tmp + sqrt(tmp * tmp - 4. * Q * beta_w));
      zplus /= 2. * beta_w;
      bzplus = (tmp + sqrt(tmp * tmp - 4. * Q * beta_w)) / 2.;
      zminus = (tmp - sqrt(tmp * tmp - 4. * Q * beta_w));
      zminus /= 2. * beta_w;
      if (fabs(zminus) < 1.e-15) zminus = 0.0;
    } else {
      zplus = tmp / beta_w - (Q / tmp) * (1. - 2. * beta_w / (tmp * tmp));
      bzplus = tmp - (Q * beta_w / tmp) * (1. - 2. * beta_w / (tmp * tmp));
      zminus = (Q / tmp) * (1. - 2. * beta_w / (tmp * tmp));
    }
    stable = 2;
    if (roots) {
      const double ome2 = (1. - E) * (1. + E);
      const double aeml = a * E - Lz;
      const double cof[5] = {-a * a * Q / ome2, 2. * (Q + aeml * aeml) / ome2, -(a * a * ome2 + Lz * Lz + Q) / ome2, 2. / ome2, -1.};
      if (QuarticRoots<CX>::solve(cof, r)) return -2;
      if (fabs(r[0] - ap) < 1e-7 && fabs(r[1] - peri) < 1e-7) {
        stable = 1;
        if (fabs(r[1] - r[2]) < 1e-7) stable = 0;
      }
      if (fabs(r[1] - peri) > 1e-7) stable = -1;
      if (peri < Kerr::rplus(a)) stable = -1;
    }
    return 0;
  }

  // Fundamental frequencies (r, theta, phi) in cycles per unit M.
  AAK_HD void frequencies(double *Om) const {
    const double r3 = r[2], r4 = r[3];
    const double rk = sqrt((ap - peri) * (r3 - r4) / ((ap - r3) * (peri - r4)));
    const double thk = sqrt(zminus / zplus);
    const double fact = p / ((1. - ecc * ecc) * sqrt(1. - E * E));
    const double rtfact = sqrt((ap - r3) * (peri - r4));
    const double art = sqrt(1. - a * a), rH = 1. + art, rC = 1. - art;
    const double ElK = Carlson::K(thk), ElEth = Carlson::E(thk), ElE = Carlson::E(rk);
    const double PiR4 = (ap - r4) * Carlson::Pi(-(peri - ap) / (peri - r4), rk) / rtfact;
    const double PiRH = (ap - r4) * Carlson::Pi(-(ap - peri) * (rH - r4) / ((peri - r4) * (ap - rH)), rk) /
                        ((ap - rH) * (rH - r4) * rtfact * art);
    const double PiRC = (ap - r4) * Carlson::Pi(-(ap - peri) * (rC - r4) / ((peri - r4) * (ap - rC)), rk) /
                        ((ap - rC) * (rC - r4) * rtfact * art);
    const double DR4 = 1. / (r4 * r4 - 2. * r4 + a * a);
    const double X = 2. * fact * Carlson::K(rk) / rtfact;
    const double Y = fact * ((ap + peri + r3 + r4) * PiR4 + rtfact * ElE) + (r4 * r4 - 0.5 * (ap - r4) * (peri - r4)) * X;
    const double Z = fact * ((2. * E * rH - a * Lz) * a * PiRH + a * (a * Lz - 2. * E * rC) * PiRC) +
                     (Lz * r4 + 2. * a * E - 2. * Lz) * r4 * X * DR4;
    const double WI = 8. * E - 2. * a * Lz, WII = 4. * E * a * a;
    const double W = E * Y + (2. * E * r4 + 4. * E + (WI * r4 - WII) * DR4) * X +
                     fact * (4. * E * PiR4 + (WI * rH - WII) * PiRH + (WII - WI * rC) * PiRC);
    const double lambda = (Y + a * a * zplus * X) * ElK - a * a * zplus * X * Carlson::E(thk);
    const double gamma = ((W + a * a * zplus * E * X) * ElK - a * a * zplus * E * X * ElEth) / lambda;
    Om[0] = p * ElK / (2. * gamma * (1. - ecc * ecc) * lambda);
    Om[1] = sqrt(bzplus) * X / (4. * gamma * lambda);
    Om[2] = ((Z - Lz * X) * ElK + Lz * X * Carlson::Pi(-zminus, thk)) / (2. * kPi * gamma * lambda);
  }
};

// Circular orbit at fixed cos(iota) (CKG.cc:9-74,116-150): E, Lz, Q, L/cos(iota).
struct CircularOrbit {
  double E, Lz, Q, Lovercosi;
  AAK_HD static double efunc(double r, double a, double Lz, double D) {
    const double rm1 = r - 1.;
    const double num = a * a * Lz * Lz * rm1 + r * D * D;
    double den = r * r * r * r * (r - 3.);
    den += a * a * r * (Lz * Lz + 2. * r * rm1);
    den += a * a * a * a * (1. + r);
    den = D * sqrt(r * den);
    den += a * Lz * (r * r - a * a);
    return num / den;
  }
  AAK_HD static double lzfunc(double r, double a, double E, double D) {
    const double tmp1 = r * (1. + r * (E * E - 1.));
    return (E * (r * r - a * a) - D * sqrt(tmp1)) / (a * (r - 1.));
  }
  AAK_HD static double qfunc(double r, double a, double E, double Lz, double D) {
    const double b = E * (a * a + r * r) - a * Lz;
    const double Q1 = b * b / D;
    const double Q2 = r * r + a * E * (a * E - 2. * Lz) + Lz * Lz;
    if (fabs(Q1 - Q2) < 1.e-9) return 0.;
    return Q1 - Q2;
  }
  AAK_HD int init(double ci, double r, double a) {
    const double D = Kerr::delta(r, a);
    E = Lz = Q = Lovercosi = 0.0;
    if (ci == 1.) {
      E = Kerr::e_pro(r, a);
    } else if (ci == -1.) {
      E = Kerr::e_ret(r, a);
    } else if (ci < 1.e-9 && ci > -1.e-9) {
      ci = 0.;
      E = efunc(r, a, 0.0, D);
      Q = qfunc(r, a, E, 0.0, D);
      Lovercosi = sqrt(Q);
    } else {
      // Bisection on E (E = (hi + lo) / 2 until |cos iota(E) - ci| < 1e-9),
      // two levels per round: the midpoint and both of its possible
      // successors are evaluated side by side (independent dependency
      // chains), then the two decisions are applied in order, so every
      // midpoint, comparison and exit is the one-level loop's.
      double lo = Kerr::e_pro(r, a);
      double hi = (r >= Kerr::isco_ret(a)) ? Kerr::e_ret(r, a) : 1.;
      auto cosi_of = [&](double e) {
        const double lz = lzfunc(r, a, e, D);
        const double q = qfunc(r, a, e, lz, D);
        return lz / sqrt(lz * lz + q);
      };
      for (int it = 0;; it += 2) {
        const double E0 = 0.5 * (hi + lo);
        const double Ea = 0.5 * (hi + E0), Eb = 0.5 * (E0 + lo);
        const double c0 = cosi_of(E0), ca = cosi_of(Ea), cb = cosi_of(Eb);
        E = E0;
        if (c0 > ci) lo = E0;
        else hi = E0;
        if (fabs(c0 - ci) < 1.e-9) break;
        if (it > 4000) return -1;
        const bool up = c0 > ci;  // lo moved: the next midpoint is Ea
        E = up ? Ea : Eb;
        const double c1 = up ? ca : cb;
        if (c1 > ci) lo = E;
        else hi = E;
        if (fabs(c1 - ci) < 1.e-9) break;
        if (it + 1 > 4000) return -1;
      }
    }
    Lz = lzfunc(r, a, E, D);
    if (ci != 0.0) {
      Q = Lz * Lz * (1. - ci * ci) / (ci * ci);
      Lovercosi = sqrt(Lz * Lz + Q);
    }
    return 0;
  }
};

// Qdot and rdot of a circular orbit from E and Lz fluxes (RRGW.cc:22-96);
// only reached when the eccentricity is exactly zero.
AAK_HD void circular_rdot(double r, double a, double Q, double E, double Lz, double Edot, double Lzdot, double *rdot) {
  const double Ep = -Edot, Lp = -Lzdot;
  const double lz2pq = Lz * Lz + Q, ome2 = 1. - E * E, lzmae = Lz - a * E, qp = Q + lzmae * lzmae;
  const double c21 = -2. * a * a * a * lzmae +
                     r * (2. * a * a * a * a * E + r * (2. * a * (lzmae - a * E) + r * (4. * a * a * E + r * (-6. * E + r * (2. * E)))));
  const double c22 = 2. * a * a * lzmae + r * (-2. * a * a * Lz + r * 2. * a * E);
  const double d = 2. * lz2pq + a * (-4 * E * Lz + a * (2. * E * E - lz2pq - a * a * ome2)) +
                   r * (2 * (a * a * (2. + ome2) + 2. * a * E * Lz - lz2pq) +
                        r * ((lz2pq - 6. - 5. * a * a * ome2) + r * (8. * ome2 + r * (-2 * ome2))));
  (void)qp;
  *rdot = -(c21 * Ep + c22 * Lp) / d;
}

// Gair-Glampedakis kludge fluxes -> dp/dt, de/dt, dcos(iota)/dt.
struct KludgeRates {
  double p, e, E, Lz, Q, a;
  double Ec, Lc, Qc, LoverCc;
  double ci, sin2i, sini, Lovercosi, ome, ope, ome2;
  double Ldotcirc, Qdotcirc;

  AAK_HD double lzdot() {
    const double pref = -6.4 * pow(ome2, 1.5) / pow(p, 3.5);
    const double t1 = 0.875 * e * e, t2 = e * e * (63. / 8. + e * e * 95. / 64), t3 = e * e * (91. / 4. + e * e * 461. / 64.);
    const double t4 = 425. * e * e / 336., t5 = 97. * e * e / 8., t6 = 302893. * e * e / 6048., t7 = 95. * e * e / 16.;
    const double c2i = ci * ci, sp = sqrt(p);
    const double c1 = -10.741956 + (28.5942157 + -9.077378144 / sp) / sp;
    const double c2 = (-1.428362761 + (10.70029768 + -33.70903016 / sp) / sp);
    const double corr = a * ((736.2086781 - 283.9553066 * sp + a * (-1325.1852209 + 483.266206498 * sp + a * (634.49936445 - 219.223848944 * sp))) * ci +
                             (82.07804475 - 25.82025864 * sp + a * (-904.16109275 + 301.477789146 * sp + a * (827.31891826 - 271.9659423 * sp))) * ci * c2i) /
                        (p * sp);
    const double hof =
        a * (c1 + a * a * c2) +
        ci * (-28.15174147 + (60.9607071973 + 40.99984205 / sp) / sp + a * a * ((-0.348161211 + (2.37258476 + -66.65840948 / sp) / sp) + a * a * (-0.715392387 + (3.21592568 + 5.28887649 / sp) / sp))) +
        c2i * a * (-7.6103411 + (128.87778309 + -475.4650442 / sp) / sp + a * a * (12.290783385 + (-113.1250548 + 306.11883292 / sp) / sp)) +
        ci * c2i * a * a * ((40.9258725 + (-347.2713496 + 886.50332051 / sp) / sp) + a * a * (-25.48313727 + (224.22721861 + -490.98212316 / sp) / sp)) +
        c2i * c2i * a * a * a * ((-9.006337706 + (91.17666278 + -297.001939215 / sp) / sp) + ci * a * (-0.64500047 + (-5.13591989 + 47.19818628 / sp) / sp)) + corr;
    const double circbit = ci + a / (pow(p, 1.5)) * (61. / 24. - c2i * 61. / 8.) - ci * 1247. / (336. * p) + 4. * kPi * ci / pow(p, 1.5) -
                           44711. * ci / (9072. * p * p) + a * a * ci * (-57. / 16. + 45. * c2i / 8.) / (p * p) + hof / (p * p * sp);
    Ldotcirc = -6.4 * circbit / pow(p, 3.5);
    return (pref * (ci * t1 + a / (pow(p, 1.5)) * (t2 - c2i * t3) - ci * t4 / p + kPi * t5 * ci / pow(p, 1.5) - t6 * ci / (p * p) +
                    a * a * t7 * ci / (p * p) + 0. + circbit));
  }
  AAK_HD double qdot() {
    const double c2i = ci * ci, sp = sqrt(p);
    const double corr = a * ((736.2086781 - 283.9553066 * sp + a * (-1325.1852209 + 483.266206498 * sp + a * (634.49936445 - 219.223848944 * sp))) +
                             (82.07804475 - 25.82025864 * sp + a * (-904.16109275 + 301.477789146 * sp + a * (827.31891826 - 271.9659423 * sp))) * c2i) /
                        (p * sp);
    const double idt = a * a * (247.1682656 / sp - 162.2684644 + c2i * (-267.5529723 / sp + 184.4645976) +
                                a * (-182.165263315 / sp + 152.125216225 + c2i * (254.0668915 / sp - 188.131613584))) / (p * p);
    const double hof =
        (-28.15174147 + (60.9607071973 + 40.99984205 / sp) / sp + a * a * ((-0.348161211 + (2.37258476 + -66.65840948 / sp) / sp) + a * a * (-0.715392387 + (3.21592568 + 5.28887649 / sp) / sp))) +
        ci * a * (-7.6103411 + (128.87778309 + -475.4650442 / sp) / sp + a * a * (12.290783385 + (-113.1250548 + 306.11883292 / sp) / sp)) +
        c2i * a * a * ((40.9258725 + (-347.2713496 + 886.50332051 / sp) / sp) + a * a * (-25.48313727 + (224.22721861 + -490.98212316 / sp) / sp)) +
        c2i * ci * a * a * a * ((-9.006337706 + (91.17666278 + -297.001939215 / sp) / sp) + ci * a * (-0.64500047 + (-5.13591989 + 47.19818628 / sp) / sp)) + corr;
    double circ = 1. - a / (pow(p, 1.5)) * (ci * 61. / 8.) - 1247. / (336. * p) + 4. * kPi / pow(p, 1.5) - 44711. / (9072. * p * p) +
                  a * a * (-57. / 16. + 45. * c2i / 8.) / (p * p) + hof / (p * p * sp);
    circ -= a * a * (-0.03093408 + -22.24163077 / p + 7.55265467 / (p * sp) + a * ci * (-3.33475762 / (sp) + 22.70130573 / p + -12.470005617 / (p * sp)) + idt) / (p * p);
    Qdotcirc = -12.8 * LoverCc * sin2i * circ / pow(p, 3.5);
    return -12.8 * Lovercosi * sin2i * (1. - e * e) * sqrt(1. - e * e) *
           (circ + 7. * e * e / 8. - (e * e * 425. / 336.) / p + kPi * (97. * e * e / 8.) / (p * sp) - (302893. * e * e / 6048.) / (p * p) -
            a * ci * e * e * (91. / 4. + 461. * e * e / 64.) / (p * sp) + a * a * (95. * e * e / 16.) / (p * p) + 0.) /
           pow(p, 3.5);
  }
  AAK_HD double edot() {
    const double pref = -6.4 * pow(ome2, 1.5) / pow(p, 5.);
    const double t1 = e * e * (73. / 24. + e * e * 37. / 96.);
    const double t2 = e * e * (823. / 24. + e * e * (949. / 32. + e * e * 491. / 192.));
    const double t3 = 9181. * e * e / 672., t4 = 1375. * e * e / 48., t5 = 172157 * e * e / 2592., t6 = 359. * e * e / 32.;
    const double fact = pow(ome2, 1.5);
    const double den = 4. * a * p * (Lc - a * Ec) - 2. * p * p * Ec * (p * p + a * a);
    const double dEdLz = (2. * ((p * p - 2. * p) * Lc + 2. * a * p * Ec)) / den;
    const double dEdQ = (-2. * p + p * p + a * a) / (4. * a * p * (Lc - a * Ec) - 2. * p * p * Ec * (p * p + a * a));
    (void)den;
    const double circbit = -(dEdLz * Ldotcirc + dEdQ * Qdotcirc);
    return (pref * (t1 - a * ci * t2 / (pow(p, 1.5)) - t3 / p + kPi * t4 / pow(p, 1.5) - t5 / (p * p) + a * a * t6 / (p * p) + 0.) +
            fact * circbit);
  }
  AAK_HD double denom(double r) const {
    const double lma = Lz - a * E, omE2 = (1. - E) * (1. + E);
    return (Q + lma * lma) + r * (-(a * a * omE2 + Lz * Lz + Q) + r * (3. + r * (-2. * omE2)));
  }
  AAK_HD double numer1(double r) const { return r * (-2. * a * (Lz - a * E) + r * (a * a * E + r * r * E)); }
  AAK_HD double numer2(double r) const {
    const double t = Lz * Lz + Q;
    return -a * a * Q + r * (2. * (t - a * E * Lz) + r * (-t));
  }
  AAK_HD double numer3(double r) const { return Lz * Lz * sini * (r * (2. - r) - a * a) / (ci * ci * ci); }

  // out = {pdot, edot, dcos(iota)/dt}
  AAK_HD int eval(double p_, double e_, double E_, double Lz_, double Q_, double a_, double *out) {
    p = p_; e = e_; E = E_; Lz = Lz_; Q = Q_; a = a_;
    ome = 1. - e;
    ope = 1. + e;
    ome2 = ome * ope;
    const double ra = p / ome, rp = p / ope;
    Lovercosi = sqrt(Lz * Lz + Q);
    ci = Lz / Lovercosi;
    sin2i = Q / (Lz * Lz + Q);
    sini = sqrt(sin2i);
    if (a != 0.0) {
      CircularOrbit c;
      if (c.init(ci, p, a)) return -1;
      Qc = c.Q; Ec = c.E; Lc = c.Lz; LoverCc = c.Lovercosi;
    } else {
      Ec = sqrt(1. - (p - 4.) / (p * (p - 3.)));
      LoverCc = p / sqrt(p - 3.);
      Qc = LoverCc * LoverCc * sin2i;
      Lc = LoverCc * ci;
    }
    const double Lzd = lzdot(), Qd = qdot(), Ed = edot();
    const double idot = (Q != 0.0) ? 0.5 * sini * ci * Qd / Q - sin2i * Lzd / sqrt(Q) : 0.;
    out[2] = -1. * sini * idot;
    if (e == 0.0) {
      circular_rdot(p, a, Q, E, Lz, -Ed, -Lzd, &out[0]);
      out[1] = 0.0;
    } else {
      const double radot = -(numer1(ra) * Ed / denom(ra) + numer2(ra) * Lzd / (Lz * denom(ra)) + numer3(ra) * idot / denom(ra));
      const double rpdot = -(numer1(rp) * Ed / denom(rp) + numer2(rp) * Lzd / (Lz * denom(rp)) + numer3(rp) * idot / denom(rp));
      out[0] = 0.5 * (radot * ome * ome + rpdot * ope * ope);
      out[1] = 0.5 * (radot * ome * ome2 - rpdot * ope * ome2) / p;
    }
    return 0;
  }
};

// (dp/dt, de/dt, dcos(iota)/dt) at a point of the inspiral.
template <class CX>
AAK_HD int inspiral_rates(double p, double e, double ci, double a, double *out) {
  double E, Lz, Q;
  if (e != 0.0) {
    Geodesic<CX> g;
    if (g.init(p, e, ci, a, 0)) return -1;
    E = g.E; Lz = g.Lz; Q = g.Q;
  } else {
    CircularOrbit c;
    if (c.init(ci, p, a)) return -1;
    E = c.E; Lz = c.Lz; Q = c.Q;
  }
  KludgeRates k;
  return k.eval(p, e, E, Lz, Q, a, out);
}

}  // namespace aak
