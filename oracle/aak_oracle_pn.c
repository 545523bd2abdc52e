/*
 * aak_oracle_pn.c -- PN maps, frequency map, fits, spline and Bessel
 * functions of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see aak_oracle.h).  Restates KSParMap.cc and
 * the GSL methods it links: the Powell hybrid root finder (MINPACK-1 user
 * guide, More/Garbow/Hillstrom 1980, as wrapped by gsl_multiroot hybrids),
 * the natural cubic spline (symmetric tridiagonal LDL^T solve, as GSL
 * cspline), linear least squares (as gsl_multifit_linear) and J_n(x) by
 * Miller's backward recurrence (Abramowitz & Stegun 9.12).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "aak_oracle_int.h"

static char o_errbuf[256] = "";
void o_set_error(const char *msg) {
  strncpy(o_errbuf, msg, sizeof(o_errbuf) - 1);
  o_errbuf[sizeof(o_errbuf) - 1] = 0;
}
const char *o_last_error(void) { return o_errbuf; }

/* ------------------------------------------------------------------ */
/* 3PN O(e^6) maps of Sago & Fujita 2015 (KSParMap.cc:86-209).          */
/* ------------------------------------------------------------------ */
double o_dvdt(double v, double e, double Y, double m, double M, double q) {
  const double v2 = v * v, v3 = v2 * v, v6 = v3 * v3;
  const double e2 = e * e, e4 = e2 * e2, e6 = e4 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double EG = O_EULERGAMMA, P2 = O_PI2;
  const double eq = (32. * pow(1. - e2, 1.5) * m * v6 * v3 * (1. + (7. * e2) / 8. +
       ((-5944. - 7040. * e2 + 8539. * e4) * v2) / 2688. -
       (v3 * ((-18432. - 55872. * e2 - 7056. * e4 + 49. * e6) * M_PI +
              48. * (1064. + 1516. * e2 + 475. * e4) * q * Y)) / 4608. +
       (v3 * v2 * ((-4791168. - 28113984. * e2 + 12239226. * e4 + 4005097. * e6) * M_PI -
                   144. * (139296. + 58408. * e2 + 249968. * e4 + 106707. * e6) * q * Y)) / 774144. +
       (v2 * v2 * (953325. * e6 + 8. * (34103. + 189. * q2 * (-329. + 815. * Y2)) +
                   12. * e2 * (-526955. + 126. * q2 * (-929. + 1431. * Y2)) +
                   3. * e4 * (-1232809. + 63. * q2 * (-1051. + 2997. * Y2)))) / 145152. +
       v6 * (117.72574285227559 - (1712. * EG) / 105. + (16. * P2) / 3. - (331. * q2) / 192. -
             (289. * M_PI * q * Y) / 6. + (145759. * q2 * Y2) / 1344. - (3424. * log(2.)) / 105. +
             e2 * (764.5906705662063 - (24503. * EG) / 210. + (229. * P2) / 6. + (2129. * q2) / 42. -
                   (4225. * M_PI * q * Y) / 24. + (27191. * q2 * Y2) / 224. + (1391. * log(2.)) / 30. - (78003. * log(3.)) / 280.) +
             e4 * (723.1604428881717 - (11663. * EG) / 140. + (109. * P2) / 4. - (56239. * q2) / 10752. -
                   (17113. * M_PI * q * Y) / 192. + (414439. * q2 * Y2) / 3584. - (418049. * log(2.)) / 84. + (3042117. * log(3.)) / 1120.) +
             e6 * (202.0979603432282 - (2461. * EG) / 560. + (23. * P2) / 16. - (3571. * q2) / 3584. -
                   (108577. * M_PI * q * Y) / 13824. + (41071. * q2 * Y2) / 1536. + (94138279. * log(2.)) / 2160. -
                   (42667641. * log(3.)) / 3584. - (1044921875. * log(5.)) / 96768.) -
             (107. * (256. + 1832. * e2 + 1308. * e4 + 69. * e6) * log(v)) / 1680.))) / (5. * M * M);
  return eq;
}

double o_dedt(double v, double e, double Y, double m, double M, double q) {
  const double v2 = v * v, v3 = v2 * v, v6 = v3 * v3;
  const double e2 = e * e, e4 = e2 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double EG = O_EULERGAMMA, P2 = O_PI2;
  const double eq = -(e * pow(1. - e2, 1.5) * m * v6 * v2 * (2451456. + 975744. * e2 +
        144. * (-54792. - 18600. * e2 + 22579. * e4) * v2 +
        84. * v3 * ((189120. + 286512. * e2 + 24217. * e4) * M_PI -
                    48. * (7032. + 5592. * e2 + 1313. * e4) * q * Y) +
        3. * v3 * v2 * ((-16885824. - 42474444. * e2 + 5678971. * e4) * M_PI -
                        48. * (106568. + 187862. * e2 + 442811. * e4) * q * Y) +
        8. * v2 * v2 * (6. * e2 * (-2070667. + 126. * q2 * (-2975. + 4009. * Y2)) +
                        8. * (-286397. + 63. * q2 * (-3179. + 5869. * Y2)) +
                        e4 * (-3506201. + 63. * q2 * (-3191. + 9009. * Y2))) +
        2451456. * v6 * (241.02948298890263 - (82283. * EG) / 1995. + (769. * P2) / 57. +
                         (180255. * q2) / 8512. - (11809. * M_PI * q * Y) / 152. + (598987. * q2 * Y2) / 8512. - (11021. * log(2.)) / 285. -
                         (234009. * log(3.)) / 5320. + e2 * (1048.5728116817568 - (297674. * EG) / 1995. + (2782. * P2) / 57. +
                                                           (536653. * q2) / 8512. - (91375. * M_PI * q * Y) / 608. + (356845. * q2 * Y2) / 8512. -
                                                           (2982946. * log(2.)) / 1995. + (1638063. * log(3.)) / 3040.) +
                         e4 * (725.9069951397873 - (1147147. * EG) / 15960. + (10721. * P2) / 456. +
                               (56509. * q2) / 9728. - (1739605. * M_PI * q * Y) / 29184. + (3248951. * q2 * Y2) / 68096. +
                               (760314287. * log(2.)) / 47880. - (1022385321. * log(3.)) / 340480. - (1044921875. * log(5.)) / 204228.) -
                         (107. * (6152. + 22256. * e2 + 10721. * e4) * log(v)) / 15960.))) / (120960. * M * M);
  return eq;
}

double o_dtdm(double v, double e, double Y, double q) {
  const double v2 = v * v, v3 = v2 * v, v4 = v2 * v2;
  const double e2 = e * e, e4 = e2 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double eq = (2. * e4 * (240. + v2 * (-120. + v * (42. * (-27. + 4. * q2) * v + (-6567. + 1996. * q2) * v3 +
                                                          48. * q * (8. + 77. * v2) * Y - 4. * q2 * v * (90. + 1577. * v2) * Y2))) +
                     e4 * e2 * (560. + v2 * (-360. + 960. * q * v * Y + 8816. * q * v3 * Y +
                                            v4 * (-15565. + 24. * q2 * (200. - 629. * Y2)) +
                                            v2 * (-2742. + 80. * q2 * (5. - 11. * Y2)))) -
                     8. * e2 * (-48. + v2 * (8. - 64. * q * v * Y - 688. * q * v3 * Y +
                                             2. * v2 * (99. + 16. * q2 * (-1. + 2. * Y2)) + v4 * (1233. + 8. * q2 * (-47. + 150. * Y2))))
                     + 16. * (16. + v2 * (24. + v2 * (27. * (2. + 5. * v2) - 48. * q * v * Y +
                                                      4. * q2 * (2. - v2 + (-2. + 3. * v2) * Y2))))) / (256. * v4);
  return eq;
}

double o_drdm(double v, double e, double Y, double q) {
  const double v2 = v * v, v3 = v2 * v;
  const double e2 = e * e, e4 = e2 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double eq = (16. + 8. * (-3. + e2) * v2 - 16. * (-3. + e2) * q * v3 * Y +
                     8. * (33. + 4. * e2 - 3. * e4) * q * v3 * v2 * Y +
                     v3 * v3 * (-351. + 132. * q2 + e2 * (-135. + 21. * e2 + 5. * e4 + 2. * (7. + e2) * q2) +
                                2. * (-204. + 13. * e2 * (-3. + e2)) * q2 * Y2) +
                     2. * v2 * v2 * (-45. + 3. * e4 + 4. * q2 * (1. - 4. * Y2) + 2. * e2 * q2 * (1. + Y2))) / (16. * v);
  return eq;
}

double o_dthetadm(double v, double e, double Y, double q) {
  const double v2 = v * v, v3 = v2 * v;
  const double e2 = e * e, e4 = e2 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double eq = (16. + 8. * (3. + e2) * v2 - 16. * (3. + e2) * q * v3 * Y -
                     8. * (3. + e2) * (5. + 3. * e2) * q * v3 * v2 * Y +
                     v3 * v3 * (135. - 54. * q2 + e2 * (5. * (27. + 9. * e2 + e4) + 2. * (-38. + e2) * q2) +
                                2. * (57. + 90. * e2 + 13. * e4) * q2 * Y2) +
                     2. * v2 * v2 * (27. + 3. * e4 + 2. * q2 * (-1. + 7. * Y2) + 2. * e2 * (9. + q2 * (1. + Y2)))) / (16. * v);
  return eq;
}

double o_dphidm(double v, double e, double Y, double q) {
  const double v2 = v * v, v3 = v2 * v;
  const double e2 = e * e, e4 = e2 * e2;
  const double Y2 = Y * Y, q2 = q * q;
  const double eq = (16. + 8. * (3. + e2) * v2 - 16. * q * v3 * (-2. + (3. + e2) * Y) -
                     8. * q * v3 * v2 * (-6. + 15. * Y + 3. * e4 * Y + 2. * e2 * (-4. + 7. * Y)) +
                     2. * v2 * v2 * (27. + 3. * e4 + 2. * q2 * (-1. + Y) * (1. + 7. * Y) + 2. * e2 * (9. + q2 * (1. + Y2))) +
                     v3 * v3 * (5. * e4 * e2 + e4 * (45. + q2 * (2. + 26. * Y2)) +
                                e2 * (135. + 4. * q2 * (-19. + 5. * Y * (-7. + 9. * Y))) + 3. * (45. + 2. * q2 * (-9. + Y * (-6. + 19. * Y))))) / (16. * v);
  return eq;
}

/* |Omega_phi| for prograde/retrograde orbits (AAK.cc:15-23). */
double o_omegaphi(double v, double e, double cosiota, double s, double M) {
  if (cosiota > 0) return o_dphidm(v, e, cosiota, s) / o_dtdm(v, e, cosiota, s) / M;
  return o_dphidm(v, e, -cosiota, -s) / o_dtdm(v, e, -cosiota, -s) / M;
}

double o_parang_phi0(double e, double psi) {
  /* KSParMap.cc:25 */
  return atan2(sqrt(1. - e * e) * sin(psi), e + cos(psi)) - e * sqrt(1. - e * e) * sin(psi) / (1. + e * cos(psi));
}

/* ------------------------------------------------------------------ */
/* Frequency map residual (KSParMap.cc:323-358)                         */
/* ------------------------------------------------------------------ */
void o_sol_fun(const double *x, const double *Omega, double M, double e,
               double iota, double *f) {
  const double v = x[0], Mm = x[1], sm = x[2];
  double Or, Ot, Op;
  if (cos(iota) > 0) {
    Or = o_drdm(v, e, cos(iota), sm) / o_dtdm(v, e, cos(iota), sm) / 2. / M_PI;
    Ot = o_dthetadm(v, e, cos(iota), sm) / o_dtdm(v, e, cos(iota), sm) / 2. / M_PI;
    Op = o_dphidm(v, e, cos(iota), sm) / o_dtdm(v, e, cos(iota), sm) / 2. / M_PI;
  } else {
    Or = o_drdm(v, e, -cos(iota), -sm) / o_dtdm(v, e, -cos(iota), -sm) / 2. / M_PI;
    Ot = o_dthetadm(v, e, -cos(iota), -sm) / o_dtdm(v, e, -cos(iota), -sm) / 2. / M_PI;
    Op = -o_dphidm(v, e, -cos(iota), -sm) / o_dtdm(v, e, -cos(iota), -sm) / 2. / M_PI;
  }
  f[0] = Omega[0] * Mm - Or * M;
  f[1] = Omega[1] * Mm - Ot * M;
  f[2] = Omega[2] * Mm - Op * M;
}

/* ------------------------------------------------------------------ */
/* Powell's hybrid method (dogleg trust region, forward-difference     */
/* Jacobian, Broyden rank-one updates, Jacobian refresh after two      */
/* failed steps; variables scaled by Jacobian column norms).  For n=3  */
/* the rank-one update is realised by refactorising the updated        */
/* triangle.                                                           */
/* ------------------------------------------------------------------ */
typedef struct { const double *Om; double M, e, iota; } pm_ctx;

static double enorm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* Householder QR of a 3x3 matrix A (row-major): A = Q R. */
static void qr3(const double A[3][3], double Q[3][3], double R[3][3], double *colnorm) {
  double W[3][3];
  memcpy(W, A, sizeof(W));
  for (int j = 0; j < 3; j++) {
    double s = 0.0;
    for (int i = 0; i < 3; i++) s += A[i][j] * A[i][j];
    if (colnorm) colnorm[j] = sqrt(s);
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) Q[i][j] = (i == j) ? 1.0 : 0.0;
  for (int k = 0; k < 3; k++) {
    double nrm = 0.0;
    for (int i = k; i < 3; i++) nrm += W[i][k] * W[i][k];
    nrm = sqrt(nrm);
    if (nrm == 0.0) continue;
    const double alpha = (W[k][k] > 0) ? -nrm : nrm;
    double v[3] = {0, 0, 0};
    for (int i = k; i < 3; i++) v[i] = W[i][k];
    v[k] -= alpha;
    double vtv = 0.0;
    for (int i = k; i < 3; i++) vtv += v[i] * v[i];
    if (vtv == 0.0) continue;
    for (int j = k; j < 3; j++) {
      double s = 0.0;
      for (int i = k; i < 3; i++) s += v[i] * W[i][j];
      const double t = 2.0 * s / vtv;
      for (int i = k; i < 3; i++) W[i][j] -= t * v[i];
    }
    /* accumulate Q = Q * H */
    for (int r = 0; r < 3; r++) {
      double s = 0.0;
      for (int i = k; i < 3; i++) s += Q[r][i] * v[i];
      const double t = 2.0 * s / vtv;
      for (int i = k; i < 3; i++) Q[r][i] -= t * v[i];
    }
  }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) R[i][j] = (j >= i) ? W[i][j] : 0.0;
}

static void pm_eval(const pm_ctx *c, const double *x, double *f) { o_sol_fun(x, c->Om, c->M, c->e, c->iota, f); }

static void pm_fdjac(const pm_ctx *c, double *x, const double *f, double J[3][3]) {
  const double eps = 1.4901161193847656e-08; /* sqrt(DBL_EPSILON) */
  for (int j = 0; j < 3; j++) {
    const double xj = x[j];
    double h = eps * fabs(xj);
    if (h == 0.0) h = eps;
    x[j] = xj + h;
    double f1[3];
    pm_eval(c, x, f1);
    x[j] = xj;
    for (int i = 0; i < 3; i++) J[i][j] = (f1[i] - f[i]) / h;
  }
}

/* dogleg step for min |R p - qtb| within |D p| <= delta */
static void pm_dogleg(const double R[3][3], const double *diag, const double *qtb, double delta, double *x) {
  const double epsmch = 2.220446049250313e-16;
  for (int j = 2; j >= 0; j--) {
    double sum = 0.0;
    for (int i = j + 1; i < 3; i++) sum += R[j][i] * x[i];
    double temp = R[j][j];
    if (temp == 0.0) {
      for (int i = 0; i <= j; i++) temp = fmax(temp, fabs(R[i][j]));
      temp = epsmch * temp;
      if (temp == 0.0) temp = epsmch;
    }
    x[j] = (qtb[j] - sum) / temp;
  }
  double wa1[3] = {0, 0, 0}, wa2[3];
  for (int j = 0; j < 3; j++) wa2[j] = diag[j] * x[j];
  const double qnorm = enorm3(wa2);
  if (qnorm <= delta) return;
  for (int j = 0; j < 3; j++) {
    const double t = qtb[j];
    for (int i = j; i < 3; i++) wa1[i] += R[j][i] * t;
    wa1[j] /= diag[j];
  }
  const double gnorm = enorm3(wa1);
  double sgnorm = 0.0, alpha = delta / qnorm;
  if (gnorm != 0.0) {
    for (int j = 0; j < 3; j++) wa1[j] = (wa1[j] / gnorm) / diag[j];
    for (int j = 0; j < 3; j++) {
      double s = 0.0;
      for (int i = j; i < 3; i++) s += R[j][i] * wa1[i];
      wa2[j] = s;
    }
    const double t = enorm3(wa2);
    sgnorm = (gnorm / t) / t;
    alpha = 0.0;
    if (sgnorm < delta) {
      const double bnorm = enorm3(qtb);
      double tt = (bnorm / gnorm) * (bnorm / qnorm) * (sgnorm / delta);
      const double dq = delta / qnorm, sd = sgnorm / delta;
      tt = tt - dq * sd * sd + sqrt((tt - dq) * (tt - dq) + (1.0 - dq * dq) * (1.0 - sd * sd));
      alpha = (dq * (1.0 - sd * sd)) / tt;
    }
  }
  const double temp = (1.0 - alpha) * fmin(sgnorm, delta);
  for (int j = 0; j < 3; j++) x[j] = temp * wa1[j] + alpha * x[j];
}

int o_parmap(double *map, const double *Omega, double p, double M, double s,
             double e, double iota) {
  pm_ctx c = {Omega, M, e, iota};
  double x[3] = {1. / sqrt(p), M, s};
  double f[3], J[3][3], Q[3][3], R[3][3], diag[3], cn[3], qtf[3];
  pm_eval(&c, x, f);
  double fnorm = enorm3(f);
  pm_fdjac(&c, x, f, J);
  qr3((const double (*)[3])J, Q, R, cn);
  for (int j = 0; j < 3; j++) diag[j] = (cn[j] == 0.0) ? 1.0 : cn[j];
  double dx0[3] = {diag[0] * x[0], diag[1] * x[1], diag[2] * x[2]};
  double delta = 100.0 * enorm3(dx0);
  if (delta == 0.0) delta = 100.0;
  for (int j = 0; j < 3; j++) qtf[j] = Q[0][j] * f[0] + Q[1][j] * f[1] + Q[2][j] * f[2];
  int iter = 1, ncsuc = 0, ncfail = 0, nslow1 = 0, nslow2 = 0, jeval = 1;
  int it = 0, status = 0;
  for (;;) {
    it++;
    /* one hybrid iteration */
    double pdir[3], dxs[3], xt[3], ft[3], w[3];
    pm_dogleg((const double (*)[3])R, diag, qtf, delta, pdir);
    for (int j = 0; j < 3; j++) {
      dxs[j] = -pdir[j];
      xt[j] = x[j] + dxs[j];
    }
    const double dd[3] = {diag[0] * dxs[0], diag[1] * dxs[1], diag[2] * dxs[2]};
    const double pnorm = enorm3(dd);
    if (iter == 1) delta = fmin(delta, pnorm);
    pm_eval(&c, xt, ft);
    if (!(isfinite(ft[0]) && isfinite(ft[1]) && isfinite(ft[2]))) { status = -2; break; }
    const double fnorm1 = enorm3(ft);
    double actred = -1.0;
    if (fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
    for (int i = 0; i < 3; i++) {
      double sum = 0.0;
      for (int j = i; j < 3; j++) sum += R[i][j] * dxs[j];
      w[i] = qtf[i] + sum;
    }
    const double temp = enorm3(w);
    double prered = 0.0;
    if (temp < fnorm) prered = 1.0 - (temp / fnorm) * (temp / fnorm);
    double ratio = 0.0;
    if (prered > 0.0) ratio = actred / prered;
    if (ratio < 0.1) {
      ncsuc = 0;
      ncfail++;
      delta = 0.5 * delta;
    } else {
      ncfail = 0;
      ncsuc++;
      if (ratio >= 0.5 || ncsuc > 1) delta = fmax(delta, pnorm / 0.5);
      if (fabs(ratio - 1.0) <= 0.1) delta = pnorm / 0.5;
    }
    const int accepted = (ratio >= 1e-4);
    if (accepted) {
      for (int j = 0; j < 3; j++) { x[j] = xt[j]; f[j] = ft[j]; }
      fnorm = fnorm1;
      iter++;
    }
    nslow1++;
    if (actred >= 0.001) nslow1 = 0;
    if (jeval) nslow2++;
    if (actred >= 0.1) nslow2 = 0;
    if (nslow2 == 5) status = -3;
    if (nslow1 == 10) status = -4;
    if (!status) {
      if (ncfail == 2) {
        pm_fdjac(&c, x, f, J);
        qr3((const double (*)[3])J, Q, R, cn);
        for (int j = 0; j < 3; j++) diag[j] = fmax(diag[j], cn[j]);
        for (int j = 0; j < 3; j++) qtf[j] = Q[0][j] * f[0] + Q[1][j] * f[1] + Q[2][j] * f[2];
        jeval = 1;
      } else {
        /* Broyden update of Q R with u = (Q^T ft - w)/pnorm, v = D^2 dx / pnorm */
        double u[3], vv[3], qtft[3];
        for (int j = 0; j < 3; j++) {
          qtft[j] = Q[0][j] * ft[0] + Q[1][j] * ft[1] + Q[2][j] * ft[2];
          u[j] = (qtft[j] - w[j]) / pnorm;
          vv[j] = diag[j] * ((diag[j] * dxs[j]) / pnorm);
          if (accepted) qtf[j] = qtft[j];
        }
        double Mt[3][3], Q2[3][3], R2[3][3], Qn[3][3], qn[3];
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Mt[i][j] = R[i][j] + u[i] * vv[j];
        qr3((const double (*)[3])Mt, Q2, R2, NULL);
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) Qn[i][j] = Q[i][0] * Q2[0][j] + Q[i][1] * Q2[1][j] + Q[i][2] * Q2[2][j];
        for (int j = 0; j < 3; j++) qn[j] = Q2[0][j] * qtf[0] + Q2[1][j] * qtf[1] + Q2[2][j] * qtf[2];
        memcpy(Q, Qn, sizeof(Q));
        memcpy(R, R2, sizeof(R));
        memcpy(qtf, qn, sizeof(qtf));
        jeval = 0;
      }
    }
    if (status) break;
    /* residual test as gsl_multiroot_test_residual(f, 1e-6) */
    const double resid = fabs(f[0]) + fabs(f[1]) + fabs(f[2]);
    if (resid < 1.e-6) break;
    if (it >= 1000) break;
  }
  map[0] = x[0];
  map[1] = x[1];
  map[2] = x[2];
  return status ? status : it;
}

/* ------------------------------------------------------------------ */
/* Natural cubic spline (tridiagonal LDL^T), evaluation by bisection.  */
/* c[] holds half the second derivative at each knot.                  */
/* ------------------------------------------------------------------ */
int o_spline_natural(const double *x, const double *y, int n, double *c) {
  if (n < 3) { o_set_error("spline needs at least 3 knots"); return -1; }
  const int sys = n - 2;
  double *dg = malloc(sizeof(double) * sys), *od = malloc(sizeof(double) * sys), *g = malloc(sizeof(double) * sys);
  double *gam = malloc(sizeof(double) * sys), *al = malloc(sizeof(double) * sys), *z = malloc(sizeof(double) * sys);
  c[0] = 0.0;
  c[n - 1] = 0.0;
  for (int i = 0; i < sys; i++) {
    const double hi = x[i + 1] - x[i];
    const double hip1 = x[i + 2] - x[i + 1];
    const double ydi = y[i + 1] - y[i];
    const double ydip1 = y[i + 2] - y[i + 1];
    const double gi = (hi != 0.0) ? 1.0 / hi : 0.0;
    const double gip1 = (hip1 != 0.0) ? 1.0 / hip1 : 0.0;
    od[i] = hip1;
    dg[i] = 2.0 * (hip1 + hi);
    g[i] = 3.0 * (ydip1 * gip1 - ydi * gi);
  }
  if (sys == 1) {
    c[1] = g[0] / dg[0];
  } else {
    al[0] = dg[0];
    gam[0] = od[0] / al[0];
    for (int i = 1; i < sys - 1; i++) {
      al[i] = dg[i] - od[i - 1] * gam[i - 1];
      gam[i] = od[i] / al[i];
    }
    al[sys - 1] = dg[sys - 1] - od[sys - 2] * gam[sys - 2];
    z[0] = g[0];
    for (int i = 1; i < sys; i++) z[i] = g[i] - gam[i - 1] * z[i - 1];
    for (int i = 0; i < sys; i++) z[i] = z[i] / al[i];
    c[1 + sys - 1] = z[sys - 1];
    for (int i = sys - 2; i >= 0; i--) c[1 + i] = z[i] - gam[i] * c[1 + i + 1];
  }
  free(dg); free(od); free(g); free(gam); free(al); free(z);
  return 0;
}

static int o_interval(const double *x, int n, double xv) {
  int lo = 0, hi = n - 1;
  while (hi > lo + 1) {
    const int m = (hi + lo) / 2;
    if (x[m] > xv) hi = m;
    else lo = m;
  }
  return lo;
}

double o_spline_eval(const double *x, const double *y, const double *c, int n, double xv) {
  const int i = o_interval(x, n, xv);
  const double xhi = x[i + 1], xlo = x[i];
  const double h = xhi - xlo;
  const double dy = y[i + 1] - y[i];
  const double dl = xv - xlo;
  const double ci = c[i], cip1 = c[i + 1];
  const double b = (dy / h) - h * (cip1 + 2.0 * ci) / 3.0;
  const double d = (cip1 - ci) / (3.0 * h);
  return y[i] + dl * (b + dl * (ci + dl * d));
}

/* ------------------------------------------------------------------ */
/* y - y0 = c0 (x - x0) + c1 (x - x0)^2 by least squares               */
/* (column-scaled Householder QR of the (n-1) x 2 design matrix).       */
/* ------------------------------------------------------------------ */
int o_polyfit(double *coeff, const double *x, const double *y, int n) {
  const int m = n - 1;
  if (m < 2) { o_set_error("polyfit needs at least 3 points"); return -1; }
  double *A0 = malloc(sizeof(double) * m), *A1 = malloc(sizeof(double) * m), *b = malloc(sizeof(double) * m);
  for (int i = 0; i < m; i++) {
    const double dx = x[i + 1] - x[0];
    A0[i] = dx;
    A1[i] = pow(x[i + 1] - x[0], 2);
    b[i] = y[i + 1] - y[0];
  }
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < m; i++) { s0 += A0[i] * A0[i]; s1 += A1[i] * A1[i]; }
  s0 = sqrt(s0); s1 = sqrt(s1);
  if (s0 == 0.0) s0 = 1.0;
  if (s1 == 0.0) s1 = 1.0;
  for (int i = 0; i < m; i++) { A0[i] /= s0; A1[i] /= s1; }
  /* Householder on column 0 */
  double cols[2][2];
  double *col[2] = {A0, A1};
  for (int k = 0; k < 2; k++) {
    double nrm = 0.0;
    for (int i = k; i < m; i++) nrm += col[k][i] * col[k][i];
    nrm = sqrt(nrm);
    const double alpha = (col[k][k] > 0) ? -nrm : nrm;
    const double vk = col[k][k] - alpha;
    double vtv = vk * vk;
    for (int i = k + 1; i < m; i++) vtv += col[k][i] * col[k][i];
    if (vtv != 0.0) {
      /* apply to remaining columns and b; v = (vk, col[k][k+1..]) */
      for (int j = k + 1; j < 2; j++) {
        double s = vk * col[j][k];
        for (int i = k + 1; i < m; i++) s += col[k][i] * col[j][i];
        const double t = 2.0 * s / vtv;
        col[j][k] -= t * vk;
        for (int i = k + 1; i < m; i++) col[j][i] -= t * col[k][i];
      }
      double s = vk * b[k];
      for (int i = k + 1; i < m; i++) s += col[k][i] * b[i];
      const double t = 2.0 * s / vtv;
      b[k] -= t * vk;
      for (int i = k + 1; i < m; i++) b[i] -= t * col[k][i];
    }
    cols[k][k] = alpha;
  }
  cols[0][1] = col[1][0];
  const double z1 = b[1] / cols[1][1];
  const double z0 = (b[0] - cols[0][1] * z1) / cols[0][0];
  coeff[0] = z0 / s0;
  coeff[1] = z1 / s1;
  free(A0); free(A1); free(b);
  return 0;
}

/* ------------------------------------------------------------------ */
/* J_k(x), k = nmin..nmax, by Miller's backward recurrence normalised   */
/* with J_0 + 2 sum J_2k = 1.                                           */
/* ------------------------------------------------------------------ */
int o_bessel_jn_array(int nmin, int nmax, double x, double *out) {
  if (nmin < 0 || nmax < nmin) { o_set_error("bad Bessel order range"); return -1; }
  if (x == 0.0) {
    for (int k = nmin; k <= nmax; k++) out[k - nmin] = (k == 0) ? 1.0 : 0.0;
    return 0;
  }
  const double ax = fabs(x);
  int top = nmax + 40 + (int)ax;
  if (top & 1) top++;
  double jp1 = 0.0, j = 1.0e-300, sum = 0.0;
  double *vals = malloc(sizeof(double) * (nmax - nmin + 1));
  for (int k = top; k >= 1; k--) {
    /* j = J_k (unnormalised), jp1 = J_{k+1} */
    if (k >= nmin && k <= nmax) vals[k - nmin] = j;
    if ((k & 1) == 0) sum += 2.0 * j;
    const double jm1 = (2.0 * k / ax) * j - jp1;
    jp1 = j;
    j = jm1;
    if (fabs(j) > 1e250) {
      j *= 1e-250; jp1 *= 1e-250; sum *= 1e-250;
      for (int kk = (k - 1 > nmin ? k - 1 : nmin); kk <= nmax; kk++)
        if (kk >= k) vals[kk - nmin] *= 1e-250;
    }
  }
  /* j = J_0 */
  if (nmin == 0) vals[0] = j;
  sum += j;
  for (int k = nmin; k <= nmax; k++) {
    double v = vals[k - nmin] / sum;
    if (x < 0 && (k & 1)) v = -v;
    out[k - nmin] = v;
  }
  free(vals);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Rotation to the NK wave frame (KSParMap.cc:532-574) with the BLAS   */
/* scaled 2-norm and in-order dot products.                             */
/* ------------------------------------------------------------------ */
static double o_nrm2(const double *v) {
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < 3; i++) {
    const double x = v[i];
    if (x != 0.0) {
      const double ax = fabs(x);
      if (scale < ax) {
        ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
        scale = ax;
      } else {
        ssq += (ax / scale) * (ax / scale);
      }
    }
  }
  return scale * sqrt(ssq);
}
static double o_dot3(const double *u, const double *v) {
  double r = 0.0;
  for (int i = 0; i < 3; i++) r += u[i] * v[i];
  return r;
}
static void o_cross(const double *u, const double *v, double *w) {
  w[0] = u[1] * v[2] - u[2] * v[1];
  w[1] = u[2] * v[0] - u[0] * v[2];
  w[2] = u[0] * v[1] - u[1] * v[0];
}

void o_rotcoeff(double *rot, double iota, double theta_S, double phi_S,
                double theta_K, double phi_K, double alpha) {
  double n[3], L[3], S[3], nxL[3], nxS[3];
  n[0] = sin(theta_S) * cos(phi_S);
  n[1] = sin(theta_S) * sin(phi_S);
  n[2] = cos(theta_S);
  S[0] = sin(theta_K) * cos(phi_K);
  S[1] = sin(theta_K) * sin(phi_K);
  S[2] = cos(theta_K);
  L[0] = cos(iota) * sin(theta_K) * cos(phi_K) + sin(iota) * (sin(alpha) * sin(phi_K) - cos(alpha) * cos(theta_K) * cos(phi_K));
  L[1] = cos(iota) * sin(theta_K) * sin(phi_K) - sin(iota) * (sin(alpha) * cos(phi_K) + cos(alpha) * cos(theta_K) * sin(phi_K));
  L[2] = cos(iota) * cos(theta_K) + sin(iota) * cos(alpha) * sin(theta_K);
  o_cross(n, L, nxL);
  o_cross(n, S, nxS);
  const double norm = o_nrm2(nxL) * o_nrm2(nxS);
  const double cosrot = o_dot3(nxL, nxS) / norm;
  double sinrot = o_dot3(L, nxS);
  sinrot -= o_dot3(S, nxL);
  sinrot /= norm;
  rot[0] = 2. * cosrot * cosrot - 1.;
  rot[1] = cosrot * sinrot;
  rot[2] = -rot[1];
  rot[3] = rot[0];
}
