// aak_math.hpp -- constants, Kerr closed forms, Carlson elliptic integrals and
// the post-Newtonian maps used by the AAK waveform.  Every function is
// __host__ __device__: the host build reproduces the reference CPU AAK bit
// for bit (same operation order, no contraction); the device kernels use the
// post-Newtonian maps (dense phase rates, amplitude), the host the rest.
#pragma once

#include <hip/hip_runtime.h>
#include <cmath>

#define AAK_HD __host__ __device__ inline

namespace aak {

// Physical constants (Constants.h:4-20), same double expressions.
constexpr double kG = 6.67259e-11;
constexpr double kC = 299792458.0;
constexpr double kMsun = 1.98892e30;
constexpr double kMsunSec = kG * kMsun / (kC * kC * kC);
constexpr double kPc = 3.0856775807e16;
constexpr double kGpc = (kPc * 1.e9) / kC;
constexpr double kYear = 31558149.8;
constexpr double kAUsec = 1.4959787066e11 / kC;
// KSParMap.h:17-18 (PI2 is the reference's truncated pi^2).
constexpr double kEulerGamma = 0.5772156649015329;
constexpr double kPi2Trunc = 9.86960440109;
constexpr double kPi = 3.14159265358979323846;

AAK_HD double dmax(double a, double b) { return a > b ? a : b; }
AAK_HD bool is_nan(double x) { return x != x; }
AAK_HD bool is_finite(double x) { return (x - x) == 0.0; }
AAK_HD double dmin(double a, double b) { return a < b ? a : b; }
AAK_HD bool bits_equal(double a, double b) {
  return __builtin_bit_cast(unsigned long long, a) == __builtin_bit_cast(unsigned long long, b);
}

// ---------------------------------------------------------------- Kerr
// Globals.h:43-179
struct Kerr {
  AAK_HD static double rplus(double a) { return 1. + sqrt((1. - a) * (1. + a)); }
  AAK_HD static double delta(double r, double a) { return r * r - 2. * r + a * a; }
  AAK_HD static double e_pro(double r, double a) {
    const double v = 1. / sqrt(r);
    return (1. - v * v * (2. - a * v)) / sqrt(1. - v * v * (3. - 2. * a * v));
  }
  AAK_HD static double e_ret(double r, double a) {
    const double v = 1. / sqrt(r);
    return (1. - v * v * (2. + a * v)) / sqrt(1. - v * v * (3. + 2. * a * v));
  }
  AAK_HD static double isco_z1(double a) {
    return 1. + (pow(1. + a, 1. / 3.) + pow(1. - a, 1. / 3.)) * pow((1. + a) * (1. - a), 1. / 3.);
  }
  AAK_HD static double isco(double a, double sgn) {
    const double z1 = isco_z1(a);
    const double z2 = sqrt(3. * a * a + z1 * z1);
    return 3. + z2 + sgn * sqrt((3. - z1) * (3. + z1 + 2. * z2));
  }
  AAK_HD static double isco_pro(double a) { return isco(a, -1.0); }
  AAK_HD static double isco_ret(double a) { return isco(a, 1.0); }
};

// ---------------------------------------------------------------- Carlson
// Duplication algorithm (Carlson, Numer. Algorithms 10 (1995) 13-26) with the
// classic stopping tolerances; the reference's src/numrec routines implement
// the same recipe and the host build reproduces their bits.
struct Carlson {
  AAK_HD static double rf(double x, double y, double z) {
    const double third = 1.0 / 3.0;
    double mu, dx, dy, dz;
    for (;;) {
      const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
      const double lam = sx * (sy + sz) + sy * sz;
      x = 0.25 * (x + lam);
      y = 0.25 * (y + lam);
      z = 0.25 * (z + lam);
      mu = third * (x + y + z);
      dx = (mu - x) / mu;
      dy = (mu - y) / mu;
      dz = (mu - z) / mu;
      if (!(dmax(dmax(fabs(dx), fabs(dy)), fabs(dz)) > 0.0025)) break;
    }
    const double e2 = dx * dy - dz * dz, e3 = dx * dy * dz;
    return (1.0 + ((1.0 / 24.0) * e2 - 0.1 - (3.0 / 44.0) * e3) * e2 + (1.0 / 14.0) * e3) / sqrt(mu);
  }
  AAK_HD static double rd(double x, double y, double z) {
    const double c1 = 3.0 / 14.0, c2 = 1.0 / 6.0, c3 = 9.0 / 22.0, c4 = 3.0 / 26.0;
    const double c5 = 0.25 * c3, c6 = 1.5 * c4;
    double sum = 0.0, fac = 1.0, mu, dx, dy, dz;
    for (;;) {
      const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
      const double lam = sx * (sy + sz) + sy * sz;
      sum += fac / (sz * (z + lam));
      fac = 0.25 * fac;
      x = 0.25 * (x + lam);
      y = 0.25 * (y + lam);
      z = 0.25 * (z + lam);
      mu = 0.2 * (x + y + 3.0 * z);
      dx = (mu - x) / mu;
      dy = (mu - y) / mu;
      dz = (mu - z) / mu;
      if (!(dmax(dmax(fabs(dx), fabs(dy)), fabs(dz)) > 0.0015)) break;
    }
    const double ea = dx * dy, eb = dz * dz;
    const double ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
    return 3.0 * sum + fac * (1.0 + ed * (-c1 + c5 * ed - c6 * dz * ee) +
                              dz * (c2 * ee + dz * (-c3 * ec + dz * c4 * ea))) / (mu * sqrt(mu));
  }
  AAK_HD static double rc(double x, double y) {
    const double third = 1.0 / 3.0;
    double w, mu, s;
    if (y > 0.0) {
      w = 1.0;
    } else {
      const double xt = x - y;
      w = sqrt(x) / sqrt(xt);
      x = xt;
      y = -y;
    }
    for (;;) {
      const double lam = 2.0 * sqrt(x) * sqrt(y) + y;
      x = 0.25 * (x + lam);
      y = 0.25 * (y + lam);
      mu = third * (x + y + y);
      s = (y - mu) / mu;
      if (!(fabs(s) > 0.0012)) break;
    }
    return w * (1.0 + s * s * (0.3 + s * (1.0 / 7.0 + s * (0.375 + s * (9.0 / 22.0))))) / sqrt(mu);
  }
  AAK_HD static double rj(double x, double y, double z, double p) {
    const double c1 = 3.0 / 14.0, c2 = 1.0 / 3.0, c3 = 3.0 / 22.0, c4 = 3.0 / 26.0;
    const double c5 = 0.75 * c3, c6 = 1.5 * c4, c7 = 0.5 * c2, c8 = c3 + c3;
    double a = 0, b = 0, rcx = 0, sum = 0.0, fac = 1.0, mu, dx, dy, dz, dp;
    double xt, yt, zt, pt;
    if (p > 0.0) {
      xt = x; yt = y; zt = z; pt = p;
    } else {
      xt = dmin(dmin(x, y), z);
      zt = dmax(dmax(x, y), z);
      yt = x + y + z - xt - zt;
      a = 1.0 / (yt - p);
      b = a * (zt - yt) * (yt - xt);
      pt = yt + b;
      rcx = rc(xt * zt / yt, p * pt / yt);
    }
    for (;;) {
      const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
      const double lam = sx * (sy + sz) + sy * sz;
      const double al = pt * (sx + sy + sz) + sx * sy * sz;
      const double be = pt + lam;
      sum += fac * rc(al * al, pt * (be * be));
      fac = 0.25 * fac;
      xt = 0.25 * (xt + lam);
      yt = 0.25 * (yt + lam);
      zt = 0.25 * (zt + lam);
      pt = 0.25 * (pt + lam);
      mu = 0.2 * (xt + yt + zt + pt + pt);
      dx = (mu - xt) / mu;
      dy = (mu - yt) / mu;
      dz = (mu - zt) / mu;
      dp = (mu - pt) / mu;
      if (!(dmax(dmax(fabs(dx), fabs(dy)), dmax(fabs(dz), fabs(dp))) > 0.0015)) break;
    }
    const double ea = dx * (dy + dz) + dy * dz, eb = dx * dy * dz, ec = dp * dp;
    const double ed = ea - 3.0 * ec, ee = eb + 2.0 * dp * (ea - ec);
    double ans = 3.0 * sum + fac * (1.0 + ed * (-c1 + c5 * ed - c6 * ee) + eb * (c7 + dp * (-c8 + dp * c4)) +
                                    dp * ea * (c2 - dp * c3) - c2 * dp * ec) / (mu * sqrt(mu));
    if (p <= 0.0) ans = a * (b * ans + 3.0 * (rcx - rf(xt, yt, zt)));
    return ans;
  }
  // Complete integrals at phi = pi/2 in the Legendre forms the geodesic
  // frequencies use (sin(pi/2) == 1 and cos(pi/2)^2 enter as in the reference).
  AAK_HD static double half_pi_s() { return sin(kPi / 2.); }
  AAK_HD static double half_pi_cc() {
    const double c = cos(kPi / 2.);
    return c * c;
  }
  AAK_HD static double K(double k) {
    const double s = half_pi_s();
    return s * rf(half_pi_cc(), (1.0 - s * k) * (1.0 + s * k), 1.0);
  }
  AAK_HD static double E(double k) {
    const double s = half_pi_s(), cc = half_pi_cc();
    const double q = (1.0 - s * k) * (1.0 + s * k), sk = s * k;
    return s * (rf(cc, q, 1.0) - (sk * sk) * rd(cc, q, 1.0) / 3.0);
  }
  AAK_HD static double Pi(double n, double k) {
    const double cc = half_pi_cc(), s = half_pi_s();
    const double enss = n * s * s, q = (1.0 - s * k) * (1.0 + s * k);
    return s * (rf(cc, q, 1.0) - enss * rj(cc, q, 1.0, 1.0 + enss) / 3.0);
  }
};

// ---------------------------------------------------------------- PN maps
// Sago & Fujita (2015) 3PN O(e^6) maps.  The four frequency polynomials are
// transcribed from the reference's device copies (kernel.cu:163-230); the
// oracle transcribes the host copies (KSParMap.cc:146-209).
struct PN {
  AAK_HD static double dtdm(double v, double e, double Y, double q) {
    const double v2 = v * v, v3 = v2 * v, v4 = v2 * v2, e2 = e * e, e4 = e2 * e2, Y2 = Y * Y, q2 = q * q;
    return (2. * e4 * (240. + v2 * (-120. + v * (42. * (-27. + 4. * q2) * v + (-6567. + 1996. * q2) * v3 +
                                                 48. * q * (8. + 77. * v2) * Y - 4. * q2 * v * (90. + 1577. * v2) * Y2))) +
            e4 * e2 * (560. + v2 * (-360. + 960. * q * v * Y + 8816. * q * v3 * Y +
                                   v4 * (-15565. + 24. * q2 * (200. - 629. * Y2)) +
                                   v2 * (-2742. + 80. * q2 * (5. - 11. * Y2)))) -
            8. * e2 * (-48. + v2 * (8. - 64. * q * v * Y - 688. * q * v3 * Y +
                                    2. * v2 * (99. + 16. * q2 * (-1. + 2. * Y2)) + v4 * (1233. + 8. * q2 * (-47. + 150. * Y2)))) +
            16. * (16. + v2 * (24. + v2 * (27. * (2. + 5. * v2) - 48. * q * v * Y +
                                           4. * q2 * (2. - v2 + (-2. + 3. * v2) * Y2))))) / (256. * v4);
  }
  AAK_HD static double drdm(double v, double e, double Y, double q) {
    const double v2 = v * v, v3 = v2 * v, e2 = e * e, e4 = e2 * e2, Y2 = Y * Y, q2 = q * q;
    return (16. + 8. * (-3. + e2) * v2 - 16. * (-3. + e2) * q * v3 * Y +
            8. * (33. + 4. * e2 - 3. * e4) * q * v3 * v2 * Y +
            v3 * v3 * (-351. + 132. * q2 + e2 * (-135. + 21. * e2 + 5. * e4 + 2. * (7. + e2) * q2) +
                       2. * (-204. + 13. * e2 * (-3. + e2)) * q2 * Y2) +
            2. * v2 * v2 * (-45. + 3. * e4 + 4. * q2 * (1. - 4. * Y2) + 2. * e2 * q2 * (1. + Y2))) / (16. * v);
  }
  AAK_HD static double dthetadm(double v, double e, double Y, double q) {
    const double v2 = v * v, v3 = v2 * v, e2 = e * e, e4 = e2 * e2, Y2 = Y * Y, q2 = q * q;
    return (16. + 8. * (3. + e2) * v2 - 16. * (3. + e2) * q * v3 * Y -
            8. * (3. + e2) * (5. + 3. * e2) * q * v3 * v2 * Y +
            v3 * v3 * (135. - 54. * q2 + e2 * (5. * (27. + 9. * e2 + e4) + 2. * (-38. + e2) * q2) +
                       2. * (57. + 90. * e2 + 13. * e4) * q2 * Y2) +
            2. * v2 * v2 * (27. + 3. * e4 + 2. * q2 * (-1. + 7. * Y2) + 2. * e2 * (9. + q2 * (1. + Y2)))) / (16. * v);
  }
  AAK_HD static double dphidm(double v, double e, double Y, double q) {
    const double v2 = v * v, v3 = v2 * v, e2 = e * e, e4 = e2 * e2, Y2 = Y * Y, q2 = q * q;
    return (16. + 8. * (3. + e2) * v2 - 16. * q * v3 * (-2. + (3. + e2) * Y) -
            8. * q * v3 * v2 * (-6. + 15. * Y + 3. * e4 * Y + 2. * e2 * (-4. + 7. * Y)) +
            2. * v2 * v2 * (27. + 3. * e4 + 2. * q2 * (-1. + Y) * (1. + 7. * Y) + 2. * e2 * (9. + q2 * (1. + Y2))) +
            v3 * v3 * (5. * e4 * e2 + e4 * (45. + q2 * (2. + 26. * Y2)) +
                       e2 * (135. + 4. * q2 * (-19. + 5. * Y * (-7. + 9. * Y))) + 3. * (45. + 2. * q2 * (-9. + Y * (-6. + 19. * Y))))) /
           (16. * v);
  }
  // |Omega_phi| of the AK orbit, prograde or retrograde (gpuAAK.cc:17-26).
  AAK_HD static double omega_phi(double v, double e, double cosi, double s, double M) {
    if (cosi > 0) return dphidm(v, e, cosi, s) / dtdm(v, e, cosi, s) / M;
    return dphidm(v, e, -cosi, -s) / dtdm(v, e, -cosi, -s) / M;
  }
  // Radiation-reaction rates of v and e (KSParMap.cc:86-144).
  AAK_HD __attribute__((always_inline)) static double dvdt_p(double v, double e, double Y, double m, double M, double q, double pw, double lv) {
    const double v2 = v * v, v3 = v2 * v, v6 = v3 * v3, e2 = e * e, e4 = e2 * e2, e6 = e4 * e2;
    const double Y2 = Y * Y, q2 = q * q, EG = kEulerGamma, P2 = kPi2Trunc, PI = kPi;
    return (32. * pw * m * v6 * v3 *
            (1. + (7. * e2) / 8. + ((-5944. - 7040. * e2 + 8539. * e4) * v2) / 2688. -
             (v3 * ((-18432. - 55872. * e2 - 7056. * e4 + 49. * e6) * PI + 48. * (1064. + 1516. * e2 + 475. * e4) * q * Y)) / 4608. +
             (v3 * v2 * ((-4791168. - 28113984. * e2 + 12239226. * e4 + 4005097. * e6) * PI -
                         144. * (139296. + 58408. * e2 + 249968. * e4 + 106707. * e6) * q * Y)) / 774144. +
             (v2 * v2 * (953325. * e6 + 8. * (34103. + 189. * q2 * (-329. + 815. * Y2)) +
                         12. * e2 * (-526955. + 126. * q2 * (-929. + 1431. * Y2)) +
                         3. * e4 * (-1232809. + 63. * q2 * (-1051. + 2997. * Y2)))) / 145152. +
             v6 * (117.72574285227559 - (1712. * EG) / 105. + (16. * P2) / 3. - (331. * q2) / 192. - (289. * PI * q * Y) / 6. +
                   (145759. * q2 * Y2) / 1344. - (3424. * log(2.)) / 105. +
                   e2 * (764.5906705662063 - (24503. * EG) / 210. + (229. * P2) / 6. + (2129. * q2) / 42. - (4225. * PI * q * Y) / 24. +
                         (27191. * q2 * Y2) / 224. + (1391. * log(2.)) / 30. - (78003. * log(3.)) / 280.) +
                   e4 * (723.1604428881717 - (11663. * EG) / 140. + (109. * P2) / 4. - (56239. * q2) / 10752. -
                         (17113. * PI * q * Y) / 192. + (414439. * q2 * Y2) / 3584. - (418049. * log(2.)) / 84. +
                         (3042117. * log(3.)) / 1120.) +
                   e6 * (202.0979603432282 - (2461. * EG) / 560. + (23. * P2) / 16. - (3571. * q2) / 3584. -
                         (108577. * PI * q * Y) / 13824. + (41071. * q2 * Y2) / 1536. + (94138279. * log(2.)) / 2160. -
                         (42667641. * log(3.)) / 3584. - (1044921875. * log(5.)) / 96768.) -
                   (107. * (256. + 1832. * e2 + 1308. * e4 + 69. * e6) * lv) / 1680.))) /
           (5. * M * M);
  }
  AAK_HD __attribute__((always_inline)) static double dedt_p(double v, double e, double Y, double m, double M, double q, double pw, double lv) {
    const double v2 = v * v, v3 = v2 * v, v6 = v3 * v3, e2 = e * e, e4 = e2 * e2;
    const double Y2 = Y * Y, q2 = q * q, EG = kEulerGamma, P2 = kPi2Trunc, PI = kPi;
    return -(e * pw * m * v6 * v2 *
             (2451456. + 975744. * e2 + 144. * (-54792. - 18600. * e2 + 22579. * e4) * v2 +
              84. * v3 * ((189120. + 286512. * e2 + 24217. * e4) * PI - 48. * (7032. + 5592. * e2 + 1313. * e4) * q * Y) +
              3. * v3 * v2 * ((-16885824. - 42474444. * e2 + 5678971. * e4) * PI - 48. * (106568. + 187862. * e2 + 442811. * e4) * q * Y) +
              8. * v2 * v2 *
                  (6. * e2 * (-2070667. + 126. * q2 * (-2975. + 4009. * Y2)) + 8. * (-286397. + 63. * q2 * (-3179. + 5869. * Y2)) +
                   e4 * (-3506201. + 63. * q2 * (-3191. + 9009. * Y2))) +
              2451456. * v6 *
                  (241.02948298890263 - (82283. * EG) / 1995. + (769. * P2) / 57. + (180255. * q2) / 8512. -
                   (11809. * PI * q * Y) / 152. + (598987. * q2 * Y2) / 8512. - (11021. * log(2.)) / 285. - (234009. * log(3.)) / 5320. +
                   e2 * (1048.5728116817568 - (297674. * EG) / 1995. + (2782. * P2) / 57. + (536653. * q2) / 8512. -
                         (91375. * PI * q * Y) / 608. + (356845. * q2 * Y2) / 8512. - (2982946. * log(2.)) / 1995. +
                         (1638063. * log(3.)) / 3040.) +
                   e4 * (725.9069951397873 - (1147147. * EG) / 15960. + (10721. * P2) / 456. + (56509. * q2) / 9728. -
                         (1739605. * PI * q * Y) / 29184. + (3248951. * q2 * Y2) / 68096. + (760314287. * log(2.)) / 47880. -
                         (1022385321. * log(3.)) / 340480. - (1044921875. * log(5.)) / 204228.) -
                   (107. * (6152. + 22256. * e2 + 10721. * e4) * lv) / 15960.))) /
           (120960. * M * M);
  }
  // dv/dt and de/dt (KSParMap.cc:86-144); the _p forms take pow(1-e^2, 1.5)
  // and log(v) precomputed, so both rates of one step share them
  AAK_HD static double dvdt(double v, double e, double Y, double m, double M, double q) {
    return dvdt_p(v, e, Y, m, M, q, pow(1. - e * e, 1.5), log(v));
  }
  AAK_HD static double dedt(double v, double e, double Y, double m, double M, double q) {
    return dedt_p(v, e, Y, m, M, q, pow(1. - e * e, 1.5), log(v));
  }
  AAK_HD __attribute__((always_inline)) static void ak_rates(double v, double e, double Y, double m, double M, double q, double *edot, double *vdot) {
    const double pw = pow(1. - e * e, 1.5), lv = log(v);
    *edot = dedt_p(v, e, Y, m, M, q, pw, lv);
    *vdot = dvdt_p(v, e, Y, m, M, q, pw, lv);
  }
};

}  // namespace aak
