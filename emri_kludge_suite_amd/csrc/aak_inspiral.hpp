// aak_inspiral.hpp -- sparse inspiral in radiation-reaction time.
//
// Integrates (p, e, cos iota) with an adaptive Bulirsch-Stoer scheme
// (modified midpoint + rational extrapolation; Stoer & Bulirsch sec. 7.2.14)
// and records up to `maxsteps` points, following GKTrajFast.cc:20-130
// (LSO table), :145-318 (Eccentric), :439-457 (LSO interpolation) and
// :459-761 (step control).  The step-control bookkeeping that the reference
// keeps in function statics is per-inspiral state here and is reset at the
// first step of every output interval.
#pragma once

#include "aak_geodesic.hpp"

namespace aak {

struct InspiralPoint {
  double t, p, e, cosi;
};

template <class CX>
struct Inspiral {
  static constexpr int kMaxK = 8;
  double a;
  double plso[3][3], elso[3], clso[3];
  int first, kmax, kopt, force_reset, err;
  double epsold, xnew, dt, t;
  double acoef[kMaxK + 2];
  double alf[kMaxK + 1][kMaxK + 1];
  double tab[4][kMaxK + 1], xtab[kMaxK + 1];

  AAK_HD static int nseq(int k) { return 2 * k; }

  // LSO p on the (e, cos iota) grid {0, 0.9995} x {1, -1}.
  AAK_HD int init(double spin) {
    a = spin;
    err = 0;
    epsold = -1.0;
    xnew = 0.0;
    first = 1;
    kmax = kopt = 0;
    force_reset = 1;
    const double r_pro = Kerr::isco_pro(a), r_ret = Kerr::isco_ret(a);
    double pl, ph, pg = 0.0;
    for (int j = 1; j <= 2; j++) {
      clso[j] = 1. - ((double)(j - 1)) * 2. / ((double)(2 - 1));
      for (int i = 1; i <= 2; i++) {
        elso[i] = ((double)(i - 1)) * 0.9995 / ((double)(2 - 1));
        if (elso[i] == 0.) {
          pl = ph = (j == 1) ? r_pro : r_ret;
        } else {
          pl = pg - 0.01;
          ph = pg + 0.1;
        }
        for (int guard = 0; ph - pl > 1.e-14; guard++) {
          pg = 0.5 * (ph + pl);
          Geodesic<CX> g;
          if (g.init(pg, elso[i], clso[j], a, 1)) return -1;
          if (g.stable == 1) ph = pg;
          else if (g.stable == -1) pl = pg;
          else if (g.stable == 0) ph = pl = pg;
          if (guard > 10000) return -1;
        }
        pg = 0.5 * (ph + pl);
        plso[j][i] = pg;
      }
    }
    return 0;
  }

  AAK_HD double p_lso(double csi, double e) const {
    int ch = 2, eh = 2;
    while (ch <= 2 && clso[ch] > csi) ch++;
    while (eh <= 2 && elso[eh] < e) eh++;
    if (ch > 2) ch = 2;
    if (eh > 2) eh = 2;
    const double ef = (elso[eh] - e) / (elso[eh] - elso[eh - 1]);
    const double cf = (clso[ch] - csi) / (clso[ch] - clso[ch - 1]);
    return (cf * (ef * plso[ch - 1][eh - 1] + (1. - ef) * plso[ch - 1][eh]) +
            (1. - cf) * (ef * plso[ch][eh - 1] + (1. - ef) * plso[ch][eh]));
  }

  AAK_HD void derivs(const double *x, double *dx) {
    double o[3];
    if (inspiral_rates<CX>(x[1], x[2], x[3], a, o)) {
      err = 1;
      o[0] = o[1] = o[2] = NAN;
    }
    dx[1] = o[0];
    dx[2] = o[1];
    dx[3] = o[2];
  }

  AAK_HD int midpoint(const double *y, const double *dydx, double htot, int nstep, double *yout) {
    double ym[4], yn[4];
    const double h = htot / nstep;
    for (int i = 1; i <= 3; i++) {
      ym[i] = y[i];
      yn[i] = y[i] + h * dydx[i];
    }
    derivs(yn, yout);
    const double h2 = 2.0 * h;
    for (int n = 2; n <= nstep; n++) {
      for (int i = 1; i <= 3; i++) {
        const double s = ym[i] + h2 * yout[i];
        ym[i] = yn[i];
        yn[i] = s;
      }
      if (yn[1] - p_lso(yn[3], 0.) < -0.05 * p_lso(yn[3], 0.)) return -1;
      if (yn[2] < 0.) return -1;
      derivs(yn, yout);
    }
    for (int i = 1; i <= 3; i++) yout[i] = 0.5 * (ym[i] + yn[i] + h * yout[i]);
    return 1;
  }

  AAK_HD void extrapolate(int iest, double xest, const double *yest, double *yz, double *dy) {
    double fx[kMaxK + 2];
    xtab[iest] = xest;
    if (iest == 1) {
      for (int j = 1; j <= 3; j++) yz[j] = tab[j][1] = dy[j] = yest[j];
      return;
    }
    for (int k = 1; k < iest; k++) fx[k + 1] = xtab[iest - k] / xest;
    for (int j = 1; j <= 3; j++) {
      double v = tab[j][1], c, yy, ddy = 0.0;
      tab[j][1] = yy = c = yest[j];
      for (int k = 2; k <= iest; k++) {
        const double b1 = fx[k] * v;
        double b = b1 - c;
        if (b != 0.) {
          b = (c - v) / b;
          ddy = c * b;
          c = b1 * b;
        } else {
          ddy = v;
        }
        if (k != iest) v = tab[j][k];
        tab[j][k] = ddy;
        yy += ddy;
      }
      dy[j] = ddy;
      yz[j] = yy;
    }
  }

  // 1 ok, -1 step underflow, -2 midpoint left the physical domain
  AAK_HD int bs_step(double *y, const double *dydx, double *xx, double htry, double eps, const double *yscal,
                     double *hnext) {
    const double SAFE1 = 0.25, SAFE2 = 0.7, REDMAX = 1.0e-5, REDMIN = 0.7, SCALMX = 0.1;
    double errk[kMaxK + 1], yerr[4], ysav[4], yseq[4];
    double errmax = 0.0, red = 0.0, scale = 0.0, h, wrkmin, fact;
    int k = 0, km = 0, reduct, exitflag = 0;
    if (eps != epsold) {
      *hnext = xnew = -1.0e29;
      const double eps1 = SAFE1 * eps;
      acoef[1] = nseq(1) + 1;
      for (int q = 1; q <= kMaxK; q++) acoef[q + 1] = acoef[q] + nseq(q + 1);
      for (int iq = 2; iq <= kMaxK; iq++)
        for (int q = 1; q < iq; q++)
          alf[q][iq] = pow(eps1, (acoef[q + 1] - acoef[iq + 1]) / ((acoef[iq + 1] - acoef[1] + 1.0) * (2 * q + 1)));
      epsold = eps;
      for (kopt = 2; kopt < kMaxK; kopt++)
        if (acoef[kopt + 1] > acoef[kopt] * alf[kopt - 1][kopt]) break;
      kmax = kopt;
    }
    h = htry;
    for (int i = 1; i <= 3; i++) ysav[i] = y[i];
    if (force_reset || *xx != xnew || h != (*hnext)) {
      first = 1;
      kopt = kmax;
    }
    force_reset = 0;
    reduct = 0;
    for (;;) {
      for (k = 1; k <= kmax; k++) {
        xnew = (*xx) + h;
        if (xnew == (*xx)) return -1;
        if (midpoint(ysav, dydx, h, nseq(k), yseq) == -1) return -2;
        const double hn = h / nseq(k);
        extrapolate(k, hn * hn, yseq, y, yerr);
        if (k != 1) {
          errmax = 1.0e-30;
          for (int i = 1; i <= 3; i++) errmax = dmax(errmax, fabs(yerr[i] / yscal[i]));
          errmax /= eps;
          km = k - 1;
          errk[km] = pow(errmax / SAFE1, 1.0 / (2 * km + 1));
        }
        if (k != 1 && (k >= kopt - 1 || first)) {
          if (errmax < 1.0) {
            exitflag = 1;
            break;
          }
          if (k == kmax || k == kopt + 1) {
            red = SAFE2 / errk[km];
            break;
          } else if (k == kopt && alf[kopt - 1][kopt] < errk[km]) {
            red = 1.0 / errk[km];
            break;
          } else if (kopt == kmax && alf[km][kmax - 1] < errk[km]) {
            red = alf[km][kmax - 1] * SAFE2 / errk[km];
            break;
          } else if (alf[km][kopt] < errk[km]) {
            red = alf[km][kopt - 1] / errk[km];
            break;
          }
        }
      }
      if (exitflag) break;
      red = dmin(red, REDMIN);
      red = dmax(red, REDMAX);
      h *= red;
      reduct = 1;
    }
    *xx = xnew;
    first = 0;
    wrkmin = 1.0e35;
    for (int q = 1; q <= km; q++) {
      fact = dmax(errk[q], SCALMX);
      const double work = fact * acoef[q + 1];
      if (work < wrkmin) {
        scale = fact;
        wrkmin = work;
        kopt = q + 1;
      }
    }
    *hnext = h / scale;
    if (kopt >= k && kopt != kmax && !reduct) {
      fact = dmax(scale / alf[kopt - 1][kopt], SCALMX);
      if (acoef[kopt + 1] * fact <= wrkmin) {
        *hnext = h / fact;
        kopt++;
      }
    }
    return 1;
  }

  // 1 interval done, 0 retry with a smaller step, -1 give up, -3 error
  AAK_HD int interval(double *ystart, double x1, double x2, double eps, double h1, double hmin) {
    double y[4], dydx[4], yscal[4];
    double x = x1, hnext = 0.0;
    double h = (x2 - x1 >= 0.0) ? fabs(h1) : -fabs(h1);
    for (int i = 1; i <= 3; i++) y[i] = ystart[i];
    force_reset = 1;
    for (long n = 1; n <= 10000000L; n++) {
      derivs(y, dydx);
      if (err) return -3;
      for (int i = 1; i <= 3; i++) yscal[i] = fabs(y[i]) + fabs(dydx[i] * h) + 1.0e-30;
      if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
      const int v = bs_step(y, dydx, &x, h, eps, yscal, &hnext);
      if (err) return -3;
      if (v == -1) return -1;
      if ((x - x2) * (x2 - x1) >= 0.0) {
        for (int i = 1; i <= 3; i++) ystart[i] = y[i];
        return 1;
      }
      if (fabs(hnext) <= hmin || v == -2) return 0;
      h = hnext;
    }
    return 0;
  }

  AAK_HD int take_step(double *x, double t1) {
    const double EPS = 1.e-15 * 1000.;
    for (int guard = 0; guard < 100000; guard++) {
      const int v = interval(x, t1, t1 + dt, EPS, dt, EPS);
      if (v == -3) return -1;
      if (v == 1) return 0;
      if (v == 0) dt /= 10.;
      else if (v == -1) {
        dt = 0.0;
        return 0;
      }
    }
    return -1;
  }

  // Eccentric inspiral from (p0, e0, cos iota0); returns the number of stored
  // points (>= 0) or a negative error code.
  AAK_HD int eccentric(double p0, double e0, double ci0, double dtstart, int maxsteps, InspiralPoint *out) {
    dt = dtstart;
    double eref = e0;
    double x[4] = {0, p0, e0, ci0}, dxe[4];
    t = 0;
    int done = 0;
    if (p0 < p_lso(ci0, e0)) done = 1;
    if (e0 == 0) done = 1;
    int i = 0;
    while (!done) {
      out[i].t = t;
      out[i].p = x[1];
      out[i].e = x[2];
      out[i].cosi = x[3];
      i++;
      if (i == maxsteps) done = 1;
      const double pprev = x[1], eprev = x[2];
      if (take_step(x, t)) return -1;
      if (dt * (x[1] - pprev) > 0.) done = 1;
      if (x[2] > 0.98 && dt * (x[2] - eprev) > 0.) done = 1;
      derivs(x, dxe);
      if (err) return -1;
      while (x[1] + dt * dxe[1] < p_lso(x[3], x[2]) && fabs(dt) > 1.e-9) dt /= 2.;
      while (x[2] + dt * dxe[2] < 0 && fabs(dt) > 1.e-9) dt /= 2.;
      if (x[2] < 0.5 * eref) {
        eref = x[2];
        dt /= 2.;
      }
      if (fabs(dt) < 1.e-8 || x[2] < 1.e-6) done = 1;
      t += dt;
    }
    return i;
  }
};

}  // namespace aak
