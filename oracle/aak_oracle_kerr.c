/*
 * aak_oracle_kerr.c -- Kerr geodesic pieces of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see aak_oracle.h).  Restates the numerics of
 * IEKG.cc, IEKGMNewt.cc, CKG.cc, GKR.cc, GKG.cc, RRGW.cc and the numerical
 * methods they call (Carlson elliptic integrals, Laguerre quartic roots,
 * 2x2 Crout LU).  The floating-point expressions keep the reference's
 * operation order so that this restatement reproduces the compiled reference
 * subset in oracle/_ref bit for bit (checked in tests/test_oracle_ref.py).
 */
#define _GNU_SOURCE
#include <math.h>
#include <complex.h>
#include <stdio.h>
#include <string.h>
#include "aak_oracle_int.h"

/* ------------------------------------------------------------------ */
/* Kerr closed forms (Globals.h:43-179)                                 */
/* ------------------------------------------------------------------ */
double ok_rplus(double a) { return 1. + sqrt((1. - a) * (1. + a)); }
double ok_delta(double r, double a) { return r * r - 2. * r + a * a; }

double ok_eeqpro(double r, double a) {
  const double v = 1. / sqrt(r);
  const double num = 1. - v * v * (2. - a * v);
  const double den = 1. - v * v * (3. - 2. * a * v);
  return num / sqrt(den);
}
double ok_eeqret(double r, double a) {
  const double v = 1. / sqrt(r);
  const double num = 1. - v * v * (2. + a * v);
  const double den = 1. - v * v * (3. + 2. * a * v);
  return num / sqrt(den);
}
double ok_lzeqpro(double r, double a) {
  const double v = 1. / sqrt(r);
  const double num = 1. - a * v * v * v * (2. - a * v);
  const double den = 1. - v * v * (3. - 2. * a * v);
  return r * v * num / sqrt(den);
}
double ok_lzeqret(double r, double a) {
  const double v = 1. / sqrt(r);
  const double num = 1. + a * v * v * v * (2. + a * v);
  const double den = 1. - v * v * (3. + 2. * a * v);
  return -r * v * num / sqrt(den);
}
static double ok_isco_z1(double a) {
  return 1. + (pow(1. + a, 1. / 3.) + pow(1. - a, 1. / 3.)) * pow((1. + a) * (1. - a), 1. / 3.);
}
double ok_isco_pro(double a) {
  const double z1 = ok_isco_z1(a);
  const double z2 = sqrt(3. * a * a + z1 * z1);
  return 3. + z2 - sqrt((3. - z1) * (3. + z1 + 2. * z2));
}
double ok_isco_ret(double a) {
  const double z1 = ok_isco_z1(a);
  const double z2 = sqrt(3. * a * a + z1 * z1);
  return 3. + z2 + sqrt((3. - z1) * (3. + z1 + 2. * z2));
}

/* ------------------------------------------------------------------ */
/* Carlson symmetric integrals by the duplication theorem              */
/* (B. C. Carlson, Numer. Algorithms 10 (1995) 13-26).  The stopping    */
/* tolerances and truncated series are Carlson's; kept identical to    */
/* the reference's routines for bit parity.                             */
/* ------------------------------------------------------------------ */
static double omax2(double a, double b) { return a > b ? a : b; }
static double omin2(double a, double b) { return a < b ? a : b; }

double o_rf(double x, double y, double z) {
  const double tol = 0.0025, third = 1.0 / 3.0;
  const double c1 = 1.0 / 24.0, c2 = 0.1, c3 = 3.0 / 44.0, c4 = 1.0 / 14.0;
  double xt = x, yt = y, zt = z, mu, dx, dy, dz;
  for (;;) {
    const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
    const double lam = sx * (sy + sz) + sy * sz;
    xt = 0.25 * (xt + lam);
    yt = 0.25 * (yt + lam);
    zt = 0.25 * (zt + lam);
    mu = third * (xt + yt + zt);
    dx = (mu - xt) / mu;
    dy = (mu - yt) / mu;
    dz = (mu - zt) / mu;
    if (!(omax2(omax2(fabs(dx), fabs(dy)), fabs(dz)) > tol)) break;
  }
  const double e2 = dx * dy - dz * dz;
  const double e3 = dx * dy * dz;
  return (1.0 + (c1 * e2 - c2 - c3 * e3) * e2 + c4 * e3) / sqrt(mu);
}

double o_rd(double x, double y, double z) {
  const double tol = 0.0015;
  const double c1 = 3.0 / 14.0, c2 = 1.0 / 6.0, c3 = 9.0 / 22.0, c4 = 3.0 / 26.0;
  const double c5 = 0.25 * c3, c6 = 1.5 * c4;
  double xt = x, yt = y, zt = z, sum = 0.0, fac = 1.0, mu, dx, dy, dz;
  for (;;) {
    const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
    const double lam = sx * (sy + sz) + sy * sz;
    sum += fac / (sz * (zt + lam));
    fac = 0.25 * fac;
    xt = 0.25 * (xt + lam);
    yt = 0.25 * (yt + lam);
    zt = 0.25 * (zt + lam);
    mu = 0.2 * (xt + yt + 3.0 * zt);
    dx = (mu - xt) / mu;
    dy = (mu - yt) / mu;
    dz = (mu - zt) / mu;
    if (!(omax2(omax2(fabs(dx), fabs(dy)), fabs(dz)) > tol)) break;
  }
  const double ea = dx * dy, eb = dz * dz;
  const double ec = ea - eb, ed = ea - 6.0 * eb, ee = ed + ec + ec;
  return 3.0 * sum + fac * (1.0 + ed * (-c1 + c5 * ed - c6 * dz * ee)
                            + dz * (c2 * ee + dz * (-c3 * ec + dz * c4 * ea))) / (mu * sqrt(mu));
}

double o_rc(double x, double y) {
  const double tol = 0.0012, third = 1.0 / 3.0;
  const double c1 = 0.3, c2 = 1.0 / 7.0, c3 = 0.375, c4 = 9.0 / 22.0;
  double xt, yt, w, mu, s;
  if (y > 0.0) { xt = x; yt = y; w = 1.0; }
  else { xt = x - y; yt = -y; w = sqrt(x) / sqrt(xt); }
  for (;;) {
    const double lam = 2.0 * sqrt(xt) * sqrt(yt) + yt;
    xt = 0.25 * (xt + lam);
    yt = 0.25 * (yt + lam);
    mu = third * (xt + yt + yt);
    s = (yt - mu) / mu;
    if (!(fabs(s) > tol)) break;
  }
  return w * (1.0 + s * s * (c1 + s * (c2 + s * (c3 + s * c4)))) / sqrt(mu);
}

double o_rj(double x, double y, double z, double p) {
  const double tol = 0.0015;
  const double c1 = 3.0 / 14.0, c2 = 1.0 / 3.0, c3 = 3.0 / 22.0, c4 = 3.0 / 26.0;
  const double c5 = 0.75 * c3, c6 = 1.5 * c4, c7 = 0.5 * c2, c8 = c3 + c3;
  double xt, yt, zt, pt, a = 0, b = 0, rcx = 0, sum = 0.0, fac = 1.0;
  double mu, dx, dy, dz, dp;
  if (p > 0.0) { xt = x; yt = y; zt = z; pt = p; }
  else {
    xt = omin2(omin2(x, y), z);
    zt = omax2(omax2(x, y), z);
    yt = x + y + z - xt - zt;
    a = 1.0 / (yt - p);
    b = a * (zt - yt) * (yt - xt);
    pt = yt + b;
    const double rho = xt * zt / yt;
    const double tau = p * pt / yt;
    rcx = o_rc(rho, tau);
  }
  for (;;) {
    const double sx = sqrt(xt), sy = sqrt(yt), sz = sqrt(zt);
    const double lam = sx * (sy + sz) + sy * sz;
    const double al0 = pt * (sx + sy + sz) + sx * sy * sz;
    const double alpha = al0 * al0;
    const double be0 = pt + lam;
    const double beta = pt * (be0 * be0);
    sum += fac * o_rc(alpha, beta);
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
    if (!(omax2(omax2(fabs(dx), fabs(dy)), omax2(fabs(dz), fabs(dp))) > tol)) break;
  }
  const double ea = dx * (dy + dz) + dy * dz;
  const double eb = dx * dy * dz;
  const double ec = dp * dp;
  const double ed = ea - 3.0 * ec;
  const double ee = eb + 2.0 * dp * (ea - ec);
  double ans = 3.0 * sum + fac * (1.0 + ed * (-c1 + c5 * ed - c6 * ee) +
                                  eb * (c7 + dp * (-c8 + dp * c4))
                                  + dp * ea * (c2 - dp * c3) - c2 * dp * ec) / (mu * sqrt(mu));
  if (p <= 0.0) ans = a * (b * ans + 3.0 * (rcx - o_rf(xt, yt, zt)));
  return ans;
}

/* Incomplete integrals in the Legendre form used by the reference
 * (NR convention for Pi: +n sin^2). */
double o_ellf(double phi, double k) {
  const double s = sin(phi);
  const double cp = cos(phi);
  if (phi > M_PI / 2.)
    return (-1. * s * o_rf(cp * cp, (1.0 - s * k) * (1.0 + s * k), 1.0) +
            2. * o_rf(0., (1.0 - k) * (1.0 + k), 1.0));
  return s * o_rf(cp * cp, (1.0 - s * k) * (1.0 + s * k), 1.0);
}
double o_elle(double phi, double k) {
  const double s = sin(phi);
  const double cp = cos(phi);
  const double cc = cp * cp;
  const double q = (1.0 - s * k) * (1.0 + s * k);
  const double sk = s * k;
  return s * (o_rf(cc, q, 1.0) - (sk * sk) * o_rd(cc, q, 1.0) / 3.0);
}
double o_ellpi(double phi, double n, double k) {
  const double cp = cos(phi);
  const double cc = cp * cp;
  const double s = sin(phi);
  const double enss = n * s * s;
  const double q = (1.0 - s * k) * (1.0 + s * k);
  return s * (o_rf(cc, q, 1.0) - enss * o_rj(cc, q, 1.0, 1.0 + enss) / 3.0);
}

/* ------------------------------------------------------------------ */
/* Laguerre's method with deflation and polishing for a quartic        */
/* (Press et al. method; behaviour of NRLaguer.cc/NRzroot.cc).          */
/* ------------------------------------------------------------------ */
static int o_laguer(const double complex *a, int m, double complex *x) {
  static const double frac[16] = {0.0, 0.5, 0.25, 0.75, 0.13, 0.38, 0.62, 0.88, 1.0,
                                   0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const double epss = 1.0e-15;
  const int mt = 20, maxit = 15 * 20;
  for (int iter = 1; iter <= maxit; iter++) {
    double complex b = a[m], d = 0.0, f = 0.0;
    double err = cabs(b);
    const double abx = cabs(*x);
    for (int j = m - 1; j >= 0; j--) {
      f = (*x) * f + d;
      d = (*x) * d + b;
      b = (*x) * b + a[j];
      err = cabs(b) + abx * err;
    }
    err *= epss;
    if (cabs(b) <= err) return 0;
    const double complex g = d / b;
    const double complex g2 = g * g;
    const double complex h = g2 - (2.0 * f / b);
    const double complex sq = csqrt((double)(m - 1) * ((double)m * h - g2));
    double complex gp = g + sq;
    const double complex gm = g - sq;
    const double abp = cabs(gp), abm = cabs(gm);
    if (abp < abm) gp = gm;
    double complex dx;
    if ((abp > abm ? abp : abm) > 0.0) {
      const double complex num = CMPLX((double)(float)m, 0.0);
      dx = num / gp;
    } else {
      dx = exp(log(1 + abx)) * CMPLX((double)cosf((float)iter), (double)sinf((float)iter));
    }
    const double complex x1 = (*x) - dx;
    if (creal(*x) == creal(x1) && cimag(*x) == cimag(x1)) return 0;
    if (iter % mt) *x = x1;
    else *x = *x - frac[iter / mt] * dx;
  }
  o_set_error("too many iterations in Laguerre root finder");
  return -1;
}

static int o_zroots(const double complex *a, int m, double complex *roots, double eps) {
  double complex ad[8];
  for (int j = 0; j <= m; j++) ad[j] = a[j];
  for (int j = m; j >= 1; j--) {
    double complex x = 0.0;
    if (o_laguer(ad, j, &x)) return -1;
    if (fabs(cimag(x)) <= 2.0 * eps * fabs(creal(x))) x = CMPLX(creal(x), 0.0);
    roots[j] = x;
    double complex b = ad[j];
    for (int jj = j - 1; jj >= 0; jj--) {
      const double complex c = ad[jj];
      ad[jj] = b;
      b = x * b + c;
    }
  }
  for (int j = 1; j <= m; j++)
    if (o_laguer(a, m, &roots[j])) return -1;
  for (int j = 2; j <= m; j++) {
    const double complex x = roots[j];
    int i;
    for (i = j - 1; i >= 1; i--) {
      if (creal(roots[i]) <= creal(x)) break;
      roots[i + 1] = roots[i];
    }
    roots[i + 1] = x;
  }
  return 0;
}

int o_quartic_roots(const double *coef, double *roots_re, double *roots_im) {
  double complex a[5], r[5];
  for (int i = 0; i < 5; i++) a[i] = CMPLX(coef[i], 0.0);
  if (o_zroots(a, 4, r, 1.e-15)) return -1;
  for (int i = 0; i < 4; i++) { roots_re[i] = creal(r[i + 1]); roots_im[i] = cimag(r[i + 1]); }
  return 0;
}

/* Real parts of the roots of R(r)/(E^2-1) in descending order
 * (IEKG.cc:112-139, GKTrajFast.cc:763-796). */
static int o_rroots(double a, double E, double Lz, double Q, double *r) {
  const double ome2 = (1. - E) * (1. + E);
  double complex cof[5], rt[5];
  const double aeml = a * E - Lz;
  cof[0] = -a * a * Q / ome2;
  cof[1] = 2. * (Q + aeml * aeml) / ome2;
  cof[2] = -(a * a * ome2 + Lz * Lz + Q) / ome2;
  cof[3] = 2. / ome2;
  cof[4] = -1.;
  if (o_zroots(cof, 4, rt, 1.e-15)) return -1;
  r[0] = creal(rt[4]);
  r[1] = creal(rt[3]);
  r[2] = creal(rt[2]);
  r[3] = creal(rt[1]);
  return 0;
}

/* ------------------------------------------------------------------ */
/* Inclined eccentric Kerr geodesic (IEKG.cc:19-226, IEKGMNewt.cc)      */
/* ------------------------------------------------------------------ */
static double o_rfunc(double r, double a, double E, double Lz, double Q) {
  /* GKG.cc:8-16 */
  const double t1 = E * (r * r + a * a) - a * Lz;
  const double lma = Lz - a * E;
  const double t2 = ok_delta(r, a) * (r * r + lma * lma + Q);
  return t1 * t1 - t2;
}

/* Newton iteration on (E, Lz) with a 2x2 Crout factorisation using
 * implicit (row-scaled) partial pivoting (IEKGMNewt.cc:7-48). */
static int o_iekg_newton(o_iekg_t *g, double *x) {
  const double tolx = 1.e-14, tolf = 1.e-14;
  for (int k = 1; k <= 200; k++) {
    const double e = x[0], lz = x[1];
    const double tisqr = g->taniota * g->taniota;
    const double cisqr = g->cosiota * g->cosiota;
    const double q = lz * lz * tisqr;
    double f0 = o_rfunc(g->peri, g->a, e, lz, q);
    double f1 = o_rfunc(g->ap, g->a, e, lz, q);
    const double aelz = g->a * e - lz;
    double m[2][2];
    m[0][0] = 2. * g->peri * (2. * g->a * aelz + g->peri * (g->a * g->a * e + g->peri * g->peri * e));
    m[0][1] = -2. * g->a * g->a * lz * tisqr + 2. * g->peri * (2. * (lz * tisqr - aelz) - g->peri * lz / cisqr);
    m[1][0] = 2. * g->ap * (2. * g->a * aelz + g->ap * (g->a * g->a * e + g->ap * g->ap * e));
    m[1][1] = -2. * g->a * g->a * lz * tisqr + 2. * g->ap * (2. * (lz * tisqr - aelz) - g->ap * lz / cisqr);
    double errf = 0.0;
    errf += fabs(f0);
    errf += fabs(f1);
    if (errf <= tolf) return 0;
    double b[2] = {-f0, -f1};
    /* implicit scaling */
    double vv[2];
    for (int i = 0; i < 2; i++) {
      double big = 0.0, t;
      if ((t = fabs(m[i][0])) > big) big = t;
      if ((t = fabs(m[i][1])) > big) big = t;
      if (big == 0.0) { o_set_error("singular Jacobian in geodesic Newton"); return -1; }
      vv[i] = 1.0 / big;
    }
    /* column 0: pivot search (>= keeps the later row on ties) */
    int imax = 0;
    double big = 0.0, dum;
    if ((dum = vv[0] * fabs(m[0][0])) >= big) { big = dum; imax = 0; }
    if ((dum = vv[1] * fabs(m[1][0])) >= big) { big = dum; imax = 1; }
    int perm0 = 0;
    if (imax != 0) {
      double t0 = m[1][0]; m[1][0] = m[0][0]; m[0][0] = t0;
      double t1 = m[1][1]; m[1][1] = m[0][1]; m[0][1] = t1;
      vv[1] = vv[0];
      perm0 = 1;
    }
    if (m[0][0] == 0.0) m[0][0] = 1.0e-20;
    dum = 1.0 / m[0][0];
    m[1][0] *= dum;
    /* column 1 */
    m[1][1] = m[1][1] - m[1][0] * m[0][1];
    if (m[1][1] == 0.0) m[1][1] = 1.0e-20;
    /* forward substitution with the row permutation */
    double s0 = b[perm0];
    b[perm0] = b[0];
    b[0] = s0;
    double s1 = b[1];
    if (s0 != 0.0) s1 -= m[1][0] * b[0];
    b[1] = s1;
    /* back substitution */
    b[1] = b[1] / m[1][1];
    double s = b[0];
    s -= m[0][1] * b[1];
    b[0] = s / m[0][0];
    double errx = 0.0;
    errx += fabs(b[0]);
    x[0] += b[0];
    errx += fabs(b[1]);
    x[1] += b[1];
    if (errx <= tolx) return 0;
  }
  return 0;
}

int o_iekg_init(o_iekg_t *g, double p, double ecc, double cosi, double a, int want_roots) {
  memset(g, 0, sizeof(*g));
  g->p = p; g->ecc = ecc; g->cosiota = cosi; g->a = a;
  if (ecc == 0) { o_set_error("geodesic with zero eccentricity"); return -1; }
  g->peri = p / (1. + ecc);
  g->ap = p / (1. - ecc);
  g->tansqiota = (1. - cosi) * (1. + cosi) / (cosi * cosi);
  g->taniota = sqrt(g->tansqiota);
  /* initial guess (IEKG.cc:92-98) */
  const double semimaj = g->peri / (1. - ecc);
  double epro = ok_eeqpro(semimaj, a);
  double eret = ok_eeqret(semimaj, a);
  if (isnan(epro) || epro > 1. - 1.e-6) epro = 1. - 1.e-6;
  if (isnan(eret) || eret > 1.) eret = 1.;
  const double eg = 0.5 * (cosi + 1.) * epro + 0.5 * (1. - cosi) * eret;
  const double lg = cosi * sqrt((1. - ecc) * (1. + ecc) / (2. - 2. * eg));
  double x[2] = {eg, lg};
  if (o_iekg_newton(g, x)) return -1;
  g->E = x[0];
  g->Lz = x[1];
  g->Q = g->Lz * g->Lz * g->tansqiota;
  g->alpha_w = g->Q + g->Lz * g->Lz;
  g->beta_w = a * a * (1. - g->E * g->E);
  const double tmp = g->alpha_w + g->beta_w;
  if (g->beta_w > 1.e-15) {
    const double sq = sqrt(tmp * tmp - 4. * g->Q * g->beta_w);
    g->zedplus = (tmp + sq);
    g->zedplus /= 2. * g->beta_w;
    g->betazedplus = (tmp + sqrt(tmp * tmp - 4. * g->Q * g->beta_w)) / 2.;
    g->zedminus = (tmp - sqrt(tmp * tmp - 4. * g->Q * g->beta_w));
    g->zedminus /= 2. * g->beta_w;
    if (fabs(g->zedminus) < 1.e-15) g->zedminus = 0.0;
  } else {
    g->zedplus = tmp / g->beta_w - (g->Q / tmp) * (1. - 2. * g->beta_w / (tmp * tmp));
    g->betazedplus = tmp - (g->Q * g->beta_w / tmp) * (1. - 2. * g->beta_w / (tmp * tmp));
    g->zedminus = (g->Q / tmp) * (1. - 2. * g->beta_w / (tmp * tmp));
  }
  g->cosiota = g->Lz / sqrt(g->alpha_w);
  g->Stable = 2; /* 2 = not classified (the reference leaves it unset) */
  if (want_roots) {
    double r[4];
    if (o_rroots(a, g->E, g->Lz, g->Q, r)) return -1;
    g->r1 = r[0]; g->r2 = r[1]; g->r3 = r[2]; g->r4 = r[3];
    if (fabs(g->r1 - g->ap) < 1e-7 && fabs(g->r2 - g->peri) < 1e-7) {
      g->Stable = 1;
      if (fabs(g->r2 - g->r3) < 1e-7) g->Stable = 0;
    }
    if (fabs(g->r2 - g->peri) > 1e-7) g->Stable = -1;
    g->p3 = g->r3 * (1 - ecc);
    g->p4 = g->r4 * (1 + ecc);
    if (g->peri < ok_rplus(a)) g->Stable = -1;
  }
  return 0;
}

/* Fundamental frequencies in cycles per M (Schmidt 2002; IEKG.cc:186-226). */
void o_iekg_frequencies(const o_iekg_t *g, double *Om) {
  const double ap = g->ap, peri = g->peri, r3 = g->r3, r4 = g->r4, a = g->a;
  const double E = g->E, Lz = g->Lz, p = g->p, ecc = g->ecc;
  const double rk = sqrt((ap - peri) * (r3 - r4) / ((ap - r3) * (peri - r4)));
  const double thk = sqrt(g->zedminus / g->zedplus);
  const double fact = p / ((1. - ecc * ecc) * sqrt(1. - E * E));
  const double rtfact = sqrt((ap - r3) * (peri - r4));
  const double art = sqrt(1. - a * a);
  const double rH = 1. + art, rC = 1. - art;
  const double hp = M_PI / 2.;
  const double ElK = o_ellf(hp, thk);
  const double ElEth = o_elle(hp, thk);
  const double ElE = o_elle(hp, rk);
  const double ElPiR4 = (ap - r4) * o_ellpi(hp, -(peri - ap) / (peri - r4), rk) / rtfact;
  const double ElPiRH = (ap - r4) * o_ellpi(hp, -(ap - peri) * (rH - r4) / ((peri - r4) * (ap - rH)), rk) /
                        ((ap - rH) * (rH - r4) * rtfact * art);
  const double ElPiRC = (ap - r4) * o_ellpi(hp, -(ap - peri) * (rC - r4) / ((peri - r4) * (ap - rC)), rk) /
                        ((ap - rC) * (rC - r4) * rtfact * art);
  const double DeltaR4 = 1. / (r4 * r4 - 2. * r4 + a * a);
  const double Xint = 2. * fact * o_ellf(hp, rk) / rtfact;
  const double Yint = fact * ((ap + peri + r3 + r4) * ElPiR4 + rtfact * ElE) + (r4 * r4 - 0.5 * (ap - r4) * (peri - r4)) * Xint;
  const double Zint = fact * ((2. * E * rH - a * Lz) * a * ElPiRH + a * (a * Lz - 2. * E * rC) * ElPiRC) +
                      (Lz * r4 + 2. * a * E - 2. * Lz) * r4 * Xint * DeltaR4;
  const double WI = 8. * E - 2. * a * Lz;
  const double WII = 4. * E * a * a;
  const double Wint = E * Yint + (2. * E * r4 + 4. * E + (WI * r4 - WII) * DeltaR4) * Xint +
                      fact * (4. * E * ElPiR4 + (WI * rH - WII) * ElPiRH + (WII - WI * rC) * ElPiRC);
  const double zp = g->zedplus;
  const double lambda = (Yint + a * a * zp * Xint) * ElK - a * a * zp * Xint * o_elle(hp, thk);
  const double gamma = ((Wint + a * a * zp * E * Xint) * ElK - a * a * zp * E * Xint * ElEth) / lambda;
  Om[0] = p * ElK / (2. * gamma * (1. - ecc * ecc) * lambda);
  Om[1] = sqrt(g->betazedplus) * Xint / (4. * gamma * lambda);
  Om[2] = ((Zint - Lz * Xint) * ElK + Lz * Xint * o_ellpi(hp, -g->zedminus, thk)) / (2. * M_PI * gamma * lambda);
}

int o_iekg(double p, double e, double cosi, double a, int want_freq, double *out) {
  o_iekg_t g;
  if (o_iekg_init(&g, p, e, cosi, a, 1)) return -1;
  out[0] = g.E; out[1] = g.Lz; out[2] = g.Q; out[3] = (double)g.Stable;
  out[4] = g.zedminus; out[5] = g.zedplus; out[6] = g.betazedplus;
  out[7] = g.r1; out[8] = g.r2; out[9] = g.r3; out[10] = g.r4;
  out[11] = g.p3; out[12] = g.p4;
  if (want_freq) {
    double Om[3];
    o_iekg_frequencies(&g, Om);
    out[13] = Om[0]; out[14] = Om[1]; out[15] = Om[2];
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Circular Kerr constants at fixed cos(iota) (CKG.cc:9-74,116-150)     */
/* ------------------------------------------------------------------ */
static double ock_efunc(double r, double a, double Lz, double Delta) {
  const double rm1 = r - 1.;
  const double numer = a * a * Lz * Lz * rm1 + r * Delta * Delta;
  double denom = r * r * r * r * (r - 3.);
  denom += a * a * r * (Lz * Lz + 2. * r * rm1);
  denom += a * a * a * a * (1. + r);
  denom = Delta * sqrt(r * denom);
  denom += a * Lz * (r * r - a * a);
  return numer / denom;
}
static double ock_lzfunc(double r, double a, double E, double Delta) {
  const double rm1 = r - 1.;
  const double tmp1 = r * (1. + r * (E * E - 1.));
  const double numer = E * (r * r - a * a) - Delta * sqrt(tmp1);
  const double denom = a * rm1;
  return numer / denom;
}
static double ock_qfunc(double r, double a, double E, double Lz, double Delta) {
  const double q1b = E * (a * a + r * r) - a * Lz;
  const double Q1 = q1b * q1b / Delta;
  const double Q2 = r * r + a * E * (a * E - 2. * Lz) + Lz * Lz;
  if (fabs(Q1 - Q2) < 1.e-9) return 0.;
  return Q1 - Q2;
}

int o_ckg_init(o_ckg_t *c, double cosi, double r, double a) {
  const double eps = 1.e-9;
  const double Delta = ok_delta(r, a);
  double E = 0, Lz = 0, Q = 0, Lovercosi = 0;
  if (cosi == 1.) E = ok_eeqpro(r, a);
  else if (cosi == -1.) E = ok_eeqret(r, a);
  else if ((cosi < eps) && (cosi > -1. * eps)) {
    cosi = 0.;
    Lz = 0.;
    E = ock_efunc(r, a, Lz, Delta);
    Q = ock_qfunc(r, a, E, Lz, Delta);
    Lovercosi = sqrt(Q);
  } else {
    double Elo = ok_eeqpro(r, a), Ehi;
    if (r >= ok_isco_ret(a)) Ehi = ok_eeqret(r, a);
    else Ehi = 1.;
    for (int it = 0;; it++) {
      E = 0.5 * (Ehi + Elo);
      Lz = ock_lzfunc(r, a, E, Delta);
      Q = ock_qfunc(r, a, E, Lz, Delta);
      const double cg = Lz / sqrt(Lz * Lz + Q);
      if (cg > cosi) Elo = E;
      else Ehi = E;
      if (fabs(cg - cosi) < eps) break;
      if (it > 4000) { o_set_error("circular-orbit energy bisection did not converge"); return -1; }
    }
  }
  Lz = ock_lzfunc(r, a, E, Delta);
  if (cosi != 0.0) {
    Q = Lz * Lz * (1. - cosi * cosi) / (cosi * cosi);
    Lovercosi = sqrt(Lz * Lz + Q);
  }
  c->E = E; c->Lz = Lz; c->Q = Q; c->Lovercosi = Lovercosi; c->cosiota = cosi;
  return 0;
}

int o_ckg(double cosi, double r, double a, double *out) {
  o_ckg_t c;
  if (o_ckg_init(&c, cosi, r, a)) return -1;
  out[0] = c.E; out[1] = c.Lz; out[2] = c.Q; out[3] = c.Lovercosi;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Gair-Glampedakis kludge fluxes (GKR.cc:28-97,195-417)               */
/* ------------------------------------------------------------------ */
typedef struct {
  double p, e, E, Lz, Q, a;
  double Ecirc, Lzcirc, Qcirc, Lovercosicirc;
  double cosiota, siniotasqr, sini, Lovercosi, ome, ope, ome2, ra, rp;
  double Ldotcirc, Qdotcirc;
} o_gkr_t;

static double ogkr_qdot(o_gkr_t *g) {
  const double p = g->p, a = g->a, e = g->e, ci = g->cosiota;
  const double csqi = ci * ci;
  const double ic3a = -0.03093408, ic3b = -22.24163077, ic3c = 7.55265467;
  const double ic4a = -3.33475762, ic4b = 22.70130573, ic4c = -12.470005617;
  const double c3a = -28.15174147, c3b = 60.9607071973, c3c = 40.99984205;
  const double c4a = -0.348161211, c4b = 2.37258476, c4c = -66.65840948;
  const double c5a = -0.715392387, c5b = 3.21592568, c5c = 5.28887649;
  const double c6a = -7.6103411, c6b = 128.87778309, c6c = -475.4650442;
  const double c7a = 12.290783385, c7b = -113.1250548, c7c = 306.11883292;
  const double c8a = 40.9258725, c8b = -347.2713496, c8c = 886.50332051;
  const double c9a = -25.48313727, c9b = 224.22721861, c9c = -490.98212316;
  const double c10a = -9.006337706, c10b = 91.17666278, c10c = -297.001939215;
  const double c11a = -0.64500047, c11b = -5.13591989, c11c = 47.19818628;
  const double sp = sqrt(p);
  const double correction = a * ((736.2086781 - 283.9553066 * sp + a * (-1325.1852209 + 483.266206498 * sp + a * (634.49936445 - 219.223848944 * sp)))
                                 + (82.07804475 - 25.82025864 * sp + a * (-904.16109275 + 301.477789146 * sp + a * (827.31891826 - 271.9659423 * sp))) *
                                   csqi) / (p * sp);
  const double idtcorrection = a * a * (247.1682656 / sp - 162.2684644 + csqi * (-267.5529723 / sp + 184.4645976)
                                        + a * (-182.165263315 / sp + 152.125216225 + csqi * (254.0668915 / sp - 188.131613584))) / (p * p);
  const double hof = (c3a + (c3b + c3c / sp) / sp + a * a * ((c4a + (c4b + c4c / sp) / sp) + a * a * (c5a + (c5b + c5c / sp) / sp))) +
                     ci * a * (c6a + (c6b + c6c / sp) / sp + a * a * (c7a + (c7b + c7c / sp) / sp)) +
                     csqi * a * a * ((c8a + (c8b + c8c / sp) / sp) + a * a * (c9a + (c9b + c9c / sp) / sp)) +
                     csqi * ci * a * a * a * ((c10a + (c10b + c10c / sp) / sp) + ci * a * (c11a + (c11b + c11c / sp) / sp)) + correction;
  double circ = 1. - a / (pow(p, 1.5)) * (ci * 61. / 8.) -
                1247. / (336. * p) + 4. * M_PI / pow(p, 1.5) - 44711. / (9072. * p * p) +
                a * a * (-57. / 16. + 45. * csqi / 8.) / (p * p) + hof / (p * p * sp);
  circ -= a * a * (ic3a + ic3b / p + ic3c / (p * sp) + a * ci * (ic4a / (sp) + ic4b / p + ic4c / (p * sp))
                   + idtcorrection) / (p * p);
  g->Qdotcirc = -12.8 * g->Lovercosicirc * g->siniotasqr * circ / pow(p, 3.5);
  const double esqcorr = 0.;
  return -12.8 * g->Lovercosi * g->siniotasqr * (1. - e * e) * sqrt(1. - e * e) *
         (circ + 7. * e * e / 8. - (e * e * 425. / 336.) / p + M_PI * (97. * e * e / 8.) / (p * sp)
          - (302893. * e * e / 6048.) / (p * p) - a * ci * e * e * (91. / 4. + 461. * e * e / 64.) / (p * sp)
          + a * a * (95. * e * e / 16.) / (p * p) + esqcorr) / pow(p, 3.5);
}

static double ogkr_edot(o_gkr_t *g) {
  const double p = g->p, a = g->a, e = g->e, ci = g->cosiota;
  const double pref = -6.4 * pow(g->ome2, 1.5) / pow(p, 5.);
  const double term1 = e * e * (73. / 24. + e * e * 37. / 96.);
  const double term2 = e * e * (823. / 24. + e * e * (949. / 32. + e * e * 491. / 192.));
  const double term3 = 9181. * e * e / 672.;
  const double term4 = 1375. * e * e / 48.;
  const double term5 = 172157 * e * e / 2592.;
  const double term6 = 359. * e * e / 32.;
  const double esqcorr = 0.;
  const double fact = pow(g->ome2, 1.5);
  const double Lc = g->Lzcirc, Ec = g->Ecirc;
  const double dEdLz = (2. * ((p * p - 2. * p) * Lc + 2. * a * p * Ec)) / (4. * a * p * (Lc - a * Ec) - 2. * p * p * Ec * (p * p + a * a));
  const double dEdQ = (-2. * p + p * p + a * a) / (4. * a * p * (Lc - a * Ec) - 2. * p * p * Ec * (p * p + a * a));
  const double circbit = -(dEdLz * g->Ldotcirc + dEdQ * g->Qdotcirc);
  return (pref * (term1 - a * ci * term2 / (pow(p, 1.5)) - term3 / p + M_PI * term4 / pow(p, 1.5)
                  - term5 / (p * p) + a * a * term6 / (p * p) + esqcorr) + fact * circbit);
}

static double ogkr_lzdot(o_gkr_t *g) {
  const double p = g->p, a = g->a, e = g->e, ci = g->cosiota;
  const double pref = -6.4 * pow(g->ome2, 1.5) / pow(p, 3.5);
  const double term1 = 0.875 * e * e;
  const double term2 = e * e * (63. / 8. + e * e * 95. / 64);
  const double term3 = e * e * (91. / 4. + e * e * 461. / 64.);
  const double term4 = 425. * e * e / 336.;
  const double term5 = 97. * e * e / 8.;
  const double term6 = 302893. * e * e / 6048.;
  const double term7 = 95. * e * e / 16.;
  const double csqi = ci * ci;
  const double esqcorr = 0.;
  const double sp = sqrt(p);
  const double c1 = -10.741956 + (28.5942157 + -9.077378144 / sp) / sp;
  const double c2 = (-1.428362761 + (10.70029768 + -33.70903016 / sp) / sp);
  const double c3a = -28.15174147, c3b = 60.9607071973, c3c = 40.99984205;
  const double c4a = -0.348161211, c4b = 2.37258476, c4c = -66.65840948;
  const double c5a = -0.715392387, c5b = 3.21592568, c5c = 5.28887649;
  const double c6a = -7.6103411, c6b = 128.87778309, c6c = -475.4650442;
  const double c7a = 12.290783385, c7b = -113.1250548, c7c = 306.11883292;
  const double c8a = 40.9258725, c8b = -347.2713496, c8c = 886.50332051;
  const double c9a = -25.48313727, c9b = 224.22721861, c9c = -490.98212316;
  const double c10a = -9.006337706, c10b = 91.17666278, c10c = -297.001939215;
  const double c11a = -0.64500047, c11b = -5.13591989, c11c = 47.19818628;
  const double correction = a * ((736.2086781 - 283.9553066 * sp + a * (-1325.1852209 + 483.266206498 * sp + a * (634.49936445 - 219.223848944 * sp))) * ci
                                 + (82.07804475 - 25.82025864 * sp + a * (-904.16109275 + 301.477789146 * sp + a * (827.31891826 - 271.9659423 * sp))) *
                                   ci * csqi) / (p * sp);
  const double hof = a * (c1 + a * a * c2) + ci * (c3a + (c3b + c3c / sp) / sp + a * a * ((c4a + (c4b + c4c / sp) / sp) + a * a * (c5a + (c5b + c5c / sp) / sp))) +
                     csqi * a * (c6a + (c6b + c6c / sp) / sp + a * a * (c7a + (c7b + c7c / sp) / sp)) +
                     ci * csqi * a * a * ((c8a + (c8b + c8c / sp) / sp) + a * a * (c9a + (c9b + c9c / sp) / sp)) +
                     csqi * csqi * a * a * a * ((c10a + (c10b + c10c / sp) / sp) + ci * a * (c11a + (c11b + c11c / sp) / sp)) + correction;
  const double circbit = ci + a / (pow(p, 1.5)) * (61. / 24. - csqi * 61. / 8.) -
                         ci * 1247. / (336. * p) + 4. * M_PI * ci / pow(p, 1.5) - 44711. * ci / (9072. * p * p) +
                         a * a * ci * (-57. / 16. + 45. * csqi / 8.) / (p * p) + hof / (p * p * sp);
  g->Ldotcirc = -6.4 * circbit / pow(p, 3.5);
  return (pref * (ci * term1 + a / (pow(p, 1.5)) * (term2 - csqi * term3) -
                  ci * term4 / p + M_PI * term5 * ci / pow(p, 1.5) - term6 * ci / (p * p) +
                  a * a * term7 * ci / (p * p) + esqcorr + circbit));
}

static double ogkr_denom(const o_gkr_t *g, double r) {
  const double lzmaE = g->Lz - g->a * g->E;
  const double omE2 = (1. - g->E) * (1. + g->E);
  const double tmp1 = g->Q + lzmaE * lzmaE;
  const double tmp2 = g->a * g->a * omE2 + g->Lz * g->Lz + g->Q;
  return (tmp1 + r * (-tmp2 + r * (3. + r * (-2. * omE2))));
}
static double ogkr_numer1(const o_gkr_t *g, double r) {
  const double lzmaE = g->Lz - g->a * g->E;
  return (r * (-2. * g->a * lzmaE + r * (g->a * g->a * g->E + r * r * g->E)));
}
static double ogkr_numer2(const o_gkr_t *g, double r) {
  const double tmp = g->Lz * g->Lz + g->Q;
  return (-g->a * g->a * g->Q + r * (2. * (tmp - g->a * g->E * g->Lz) + r * (-tmp)));
}
static double ogkr_numer3(const o_gkr_t *g, double r) {
  return (g->Lz * g->Lz * g->sini * (r * (2. - r) - g->a * g->a) / (g->cosiota * g->cosiota * g->cosiota));
}

/* Qdot and rdot of a circular orbit from E and Lz fluxes (RRGW.cc:22-96). */
static void o_qdotrdot(double r, double a, double Q, double E, double Lz,
                       double Edot, double Lzdot, double *Qdot, double *rdot) {
  const double Ep = -Edot, Lp = -Lzdot;
  const double lz2pq = Lz * Lz + Q;
  const double ome2 = 1. - E * E;
  const double lzmae = Lz - a * E;
  const double qplzmae2 = Q + lzmae * lzmae;
  const double c11_6 = -4. * E * ome2, c11_5 = 12. * E;
  const double c11_4 = -2. * E * (a * a * ome2 + 3. * lz2pq);
  const double c11_3 = 8. * (a * a * E * (1. + ome2) + E * lz2pq - 2. * a * Lz);
  const double c11_2 = -2. * a * (a * E * (6. + lz2pq + a * a * ome2) - 6. * Lz);
  const double c11_1 = 4. * a * a * E * qplzmae2;
  const double c11_0 = -4. * a * lzmae * qplzmae2;
  const double c11 = c11_0 + r * (c11_1 + r * (c11_2 + r * (c11_3 + r * (c11_4 + r * (c11_5 + r * c11_6)))));
  const double c12_4 = -4. * Lz * ome2, c12_3 = 16. * ome2 * lzmae;
  const double c12_2 = 2. * (Lz * (a * a * ome2 + lz2pq) - 6. * lzmae);
  const double c12_1 = -4. * Lz * qplzmae2, c12_0 = 4. * lzmae * qplzmae2;
  const double c12 = c12_0 + r * (c12_1 + r * (c12_2 + r * (c12_3 + r * c12_4)));
  const double c21_5 = 2. * E, c21_4 = -6. * E, c21_3 = 4. * a * a * E;
  const double c21_2 = 2. * a * (lzmae - a * E), c21_1 = 2. * a * a * a * a * E;
  const double c21_0 = -2. * a * a * a * lzmae;
  const double c21 = c21_0 + r * (c21_1 + r * (c21_2 + r * (c21_3 + r * (c21_4 + r * c21_5))));
  const double c22 = 2. * a * a * lzmae + r * (-2. * a * a * Lz + r * 2. * a * E);
  const double d_4 = -2 * ome2, d_3 = 8. * ome2, d_2 = lz2pq - 6. - 5. * a * a * ome2;
  const double d_1 = 2 * (a * a * (2. + ome2) + 2. * a * E * Lz - lz2pq);
  const double d_0 = 2. * lz2pq + a * (-4 * E * Lz + a * (2. * E * E - lz2pq - a * a * ome2));
  const double d = d_0 + r * (d_1 + r * (d_2 + r * (d_3 + r * d_4)));
  *Qdot = -(c11 * Ep + c12 * Lp) / d;
  *rdot = -(c21 * Ep + c22 * Lp) / d;
}

int o_gkr(double p, double e, double E, double Lz, double Q, double a, double *out) {
  o_gkr_t g;
  memset(&g, 0, sizeof(g));
  g.p = p; g.e = e; g.E = E; g.Lz = Lz; g.Q = Q; g.a = a;
  g.ome = 1. - e;
  g.ope = 1. + e;
  g.ome2 = g.ome * g.ope;
  g.ra = p / g.ome;
  g.rp = p / g.ope;
  g.Lovercosi = sqrt(Lz * Lz + Q);
  g.cosiota = Lz / g.Lovercosi;
  g.siniotasqr = Q / (Lz * Lz + Q);
  g.sini = sqrt(g.siniotasqr);
  if (a != 0.0) {
    o_ckg_t c;
    if (o_ckg_init(&c, g.cosiota, p, a)) return -1;
    g.Qcirc = c.Q; g.Ecirc = c.E; g.Lzcirc = c.Lz; g.Lovercosicirc = c.Lovercosi;
  } else {
    g.Ecirc = sqrt(1. - (p - 4.) / (p * (p - 3.)));
    g.Lovercosicirc = p / sqrt(p - 3.);
    g.Qcirc = g.Lovercosicirc * g.Lovercosicirc * g.siniotasqr;
    g.Lzcirc = g.Lovercosicirc * g.cosiota;
  }
  const double Lzdot = ogkr_lzdot(&g);
  const double Qdot = ogkr_qdot(&g);
  const double Edot = ogkr_edot(&g);
  double idot;
  if (Q != 0.0) idot = 0.5 * g.sini * g.cosiota * Qdot / Q - g.siniotasqr * Lzdot / sqrt(Q);
  else idot = 0.;
  const double cosidot = -1. * g.sini * idot;
  double pdot, edot;
  if (e == 0.0) {
    double junk;
    o_qdotrdot(p, a, Q, E, Lz, -Edot, -Lzdot, &junk, &pdot);
    edot = 0.0;
  } else {
    const double ra = g.ra, rp = g.rp;
    const double radot = -(ogkr_numer1(&g, ra) * Edot / ogkr_denom(&g, ra) + ogkr_numer2(&g, ra) * Lzdot / (Lz * ogkr_denom(&g, ra)) + ogkr_numer3(&g, ra) * idot / ogkr_denom(&g, ra));
    const double rpdot = -(ogkr_numer1(&g, rp) * Edot / ogkr_denom(&g, rp) + ogkr_numer2(&g, rp) * Lzdot / (Lz * ogkr_denom(&g, rp)) + ogkr_numer3(&g, rp) * idot / ogkr_denom(&g, rp));
    pdot = 0.5 * (radot * g.ome * g.ome + rpdot * g.ope * g.ope);
    edot = 0.5 * (radot * g.ome * g.ome2 - rpdot * g.ope * g.ome2) / p;
  }
  out[0] = pdot; out[1] = edot; out[2] = cosidot;
  return 0;
}

int o_param_derivs(double p, double e, double cosi, double a, double *out) {
  double E, Lz, Q;
  if (e != 0.0) {
    o_iekg_t g;
    if (o_iekg_init(&g, p, e, cosi, a, 0)) return -1;
    E = g.E; Lz = g.Lz; Q = g.Q;
  } else {
    o_ckg_t c;
    if (o_ckg_init(&c, cosi, p, a)) return -1;
    E = c.E; Lz = c.Lz; Q = c.Q;
  }
  return o_gkr(p, e, E, Lz, Q, a, out);
}
