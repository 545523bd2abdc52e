/*
 * aak_oracle_traj.c -- sparse inspiral of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see aak_oracle.h).  Restates GKTrajFast.cc:
 * the LSO table (:20-130), the eccentric inspiral (:145-318), the LSO
 * interpolation (:439-457) and the Bulirsch-Stoer integrator with modified
 * midpoint steps and rational extrapolation (:459-761; Stoer & Bulirsch,
 * "Introduction to Numerical Analysis", sec. 7.2.14; Deuflhard 1983).
 *
 * The reference keeps the integrator's bookkeeping (first, kopt, xnew and
 * the work tables) in function-scope statics and passes an uninitialised
 * step-size hint to the first integrator step of every interval.  Here that
 * state lives in o_traj_t and is reset at the first step of every interval,
 * which is what the reference does whenever the uninitialised hint differs
 * from the requested step.
 */
#define _GNU_SOURCE
#include <math.h>
#include <string.h>
#include <stdio.h>
#include "aak_oracle_int.h"

#define KMAXX 8
#define IMAXX (KMAXX + 1)

typedef struct {
  double a;
  double p_lso[3][3], e_lso[3], csi_lso[3]; /* 1-based */
  /* integrator bookkeeping */
  int bs_ready, first, kmax, kopt, force_reset;
  double epsold, xnew;
  double acoef[IMAXX + 1];
  double alf[KMAXX + 1][KMAXX + 1];
  /* rational-extrapolation tableau */
  double d[4][KMAXX + 1];
  double xtab[KMAXX + 1];
  double dt, t;
  int err;
} o_traj_t;

static const int o_nseq[IMAXX + 1] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 18};

static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }

static int ot_init(o_traj_t *T, double cosi, double a) {
  memset(T, 0, sizeof(*T));
  T->a = a;
  (void)cosi;
  const double r_ipro = ok_isco_pro(a);
  const double r_iret = ok_isco_ret(a);
  double p_l, p_h, p_g = 0.0;
  for (int j = 1; j <= 2; j++) {
    T->csi_lso[j] = 1. - ((double)(j - 1)) * 2. / ((double)(2 - 1));
    for (int i = 1; i <= 2; i++) {
      T->e_lso[i] = ((double)(i - 1)) * 0.9995 / ((double)(2 - 1));
      if (T->e_lso[i] == 0.) {
        if (fabs(T->csi_lso[j] - 1.0) < 1.e-38) {
          p_l = p_h = r_ipro;
          T->csi_lso[j] = 1.0;
        } else if (fabs(T->csi_lso[j] + 1.0) < 1.e-38) {
          p_l = p_h = r_iret;
          T->csi_lso[j] = -1.0;
        } else {
          /* circular inclined LSO: unreachable with a 2-point cos(iota) grid */
          o_set_error("inclined circular LSO branch is not reachable");
          return -1;
        }
      } else {
        p_l = p_g - 0.01;
        p_h = p_g + 0.1;
      }
      int guard = 0;
      while (p_h - p_l > 1.e-14) {
        p_g = 0.5 * (p_h + p_l);
        o_iekg_t g;
        if (o_iekg_init(&g, p_g, T->e_lso[i], T->csi_lso[j], a, 1)) return -1;
        if (g.Stable == 1) p_h = p_g;
        else if (g.Stable == -1) p_l = p_g;
        else if (g.Stable == 0) p_h = p_l = p_g;
        if (++guard > 10000) { o_set_error("LSO bisection stalled"); return -1; }
      }
      p_g = 0.5 * (p_h + p_l);
      T->p_lso[j][i] = p_g;
    }
  }
  return 0;
}

static double ot_plso(const o_traj_t *T, double csi, double e) {
  int csihi = 2;
  while (csihi <= 2 && T->csi_lso[csihi] > csi) csihi++;
  int ehi = 2;
  while (ehi <= 2 && T->e_lso[ehi] < e) ehi++;
  if (csihi > 2) csihi = 2; /* outside the table: the reference reads past it */
  if (ehi > 2) ehi = 2;
  const double efrac = (T->e_lso[ehi] - e) / (T->e_lso[ehi] - T->e_lso[ehi - 1]);
  const double csifrac = (T->csi_lso[csihi] - csi) / (T->csi_lso[csihi] - T->csi_lso[csihi - 1]);
  return (csifrac * (efrac * T->p_lso[csihi - 1][ehi - 1] + (1. - efrac) * T->p_lso[csihi - 1][ehi]) +
          (1. - csifrac) * (efrac * T->p_lso[csihi][ehi - 1] + (1. - efrac) * T->p_lso[csihi][ehi]));
}

static void ot_derivs(o_traj_t *T, const double *x, double *dxdt) {
  double out[3];
  if (o_param_derivs(x[1], x[2], x[3], T->a, out)) {
    T->err = 1;
    dxdt[1] = dxdt[2] = dxdt[3] = NAN;
    return;
  }
  dxdt[1] = out[0];
  dxdt[2] = out[1];
  dxdt[3] = out[2];
}

/* modified midpoint (GKTrajFast.cc:716-761); 1 = ok, -1 = left the domain */
static int ot_mmid(o_traj_t *T, const double *y, const double *dydx, double xs,
                   double htot, int nstep, double *yout) {
  double ym[4], yn[4];
  double h = htot / nstep;
  for (int i = 1; i <= 3; i++) {
    ym[i] = y[i];
    yn[i] = y[i] + h * dydx[i];
  }
  double x = xs + h;
  (void)x;
  ot_derivs(T, yn, yout);
  const double h2 = 2.0 * h;
  for (int n = 2; n <= nstep; n++) {
    for (int i = 1; i <= 3; i++) {
      const double sw = ym[i] + h2 * yout[i];
      ym[i] = yn[i];
      yn[i] = sw;
    }
    if (yn[1] - ot_plso(T, yn[3], 0.) < -0.05 * ot_plso(T, yn[3], 0.)) return -1;
    if (yn[2] < 0.) return -1;
    x += h;
    ot_derivs(T, yn, yout);
  }
  for (int i = 1; i <= 3; i++) yout[i] = 0.5 * (ym[i] + yn[i] + h * yout[i]);
  return 1;
}

/* rational extrapolation to h -> 0 (Bulirsch-Stoer) */
static void ot_rzextr(o_traj_t *T, int iest, double xest, const double *yest,
                      double *yz, double *dy) {
  double fx[KMAXX + 2];
  T->xtab[iest] = xest;
  if (iest == 1) {
    for (int j = 1; j <= 3; j++) {
      yz[j] = yest[j];
      T->d[j][1] = yest[j];
      dy[j] = yest[j];
    }
    return;
  }
  for (int k = 1; k < iest; k++) fx[k + 1] = T->xtab[iest - k] / xest;
  for (int j = 1; j <= 3; j++) {
    double v = T->d[j][1], c, yy, ddy = 0.0;
    T->d[j][1] = yy = c = yest[j];
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
      if (k != iest) v = T->d[j][k];
      T->d[j][k] = ddy;
      yy += ddy;
    }
    dy[j] = ddy;
    yz[j] = yy;
  }
}

/* one adaptive Bulirsch-Stoer step (GKTrajFast.cc:546-672);
 * 1 ok, -1 step underflow, -2 midpoint left the domain */
static int ot_bsstep(o_traj_t *T, double *y, const double *dydx, double *xx,
                     double htry, double eps, const double *yscal, double *hdid, double *hnext) {
  const double SAFE1 = 0.25, SAFE2 = 0.7, REDMAX = 1.0e-5, REDMIN = 0.7, SCALMX = 0.1, TINY = 1.0e-30;
  double err[KMAXX + 1], yerr[4], ysav[4], yseq[4];
  double errmax = 0.0, red = 0.0, scale = 0.0, h, wrkmin, work, fact, xest;
  int k = 0, km = 0, reduct, exitflag = 0;
  if (eps != T->epsold) {
    *hnext = T->xnew = -1.0e29;
    const double eps1 = SAFE1 * eps;
    T->acoef[1] = o_nseq[1] + 1;
    for (int kk = 1; kk <= KMAXX; kk++) T->acoef[kk + 1] = T->acoef[kk] + o_nseq[kk + 1];
    for (int iq = 2; iq <= KMAXX; iq++)
      for (int kk = 1; kk < iq; kk++)
        T->alf[kk][iq] = pow(eps1, (T->acoef[kk + 1] - T->acoef[iq + 1]) /
                                       ((T->acoef[iq + 1] - T->acoef[1] + 1.0) * (2 * kk + 1)));
    T->epsold = eps;
    for (T->kopt = 2; T->kopt < KMAXX; T->kopt++)
      if (T->acoef[T->kopt + 1] > T->acoef[T->kopt] * T->alf[T->kopt - 1][T->kopt]) break;
    T->kmax = T->kopt;
  }
  h = htry;
  for (int i = 1; i <= 3; i++) ysav[i] = y[i];
  if (T->force_reset || *xx != T->xnew || h != (*hnext)) {
    T->first = 1;
    T->kopt = T->kmax;
  }
  T->force_reset = 0;
  reduct = 0;
  for (;;) {
    for (k = 1; k <= T->kmax; k++) {
      T->xnew = (*xx) + h;
      if (T->xnew == (*xx)) return -1;
      if (ot_mmid(T, ysav, dydx, *xx, h, o_nseq[k], yseq) == -1) return -2;
      const double hn = h / o_nseq[k];
      xest = hn * hn;
      ot_rzextr(T, k, xest, yseq, y, yerr);
      if (k != 1) {
        errmax = TINY;
        for (int i = 1; i <= 3; i++) errmax = dmax(errmax, fabs(yerr[i] / yscal[i]));
        errmax /= eps;
        km = k - 1;
        err[km] = pow(errmax / SAFE1, 1.0 / (2 * km + 1));
      }
      if (k != 1 && (k >= T->kopt - 1 || T->first)) {
        if (errmax < 1.0) {
          exitflag = 1;
          break;
        }
        if (k == T->kmax || k == T->kopt + 1) {
          red = SAFE2 / err[km];
          break;
        } else if (k == T->kopt && T->alf[T->kopt - 1][T->kopt] < err[km]) {
          red = 1.0 / err[km];
          break;
        } else if (T->kopt == T->kmax && T->alf[km][T->kmax - 1] < err[km]) {
          red = T->alf[km][T->kmax - 1] * SAFE2 / err[km];
          break;
        } else if (T->alf[km][T->kopt] < err[km]) {
          red = T->alf[km][T->kopt - 1] / err[km];
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
  *xx = T->xnew;
  *hdid = h;
  T->first = 0;
  wrkmin = 1.0e35;
  for (int kk = 1; kk <= km; kk++) {
    fact = dmax(err[kk], SCALMX);
    work = fact * T->acoef[kk + 1];
    if (work < wrkmin) {
      scale = fact;
      wrkmin = work;
      T->kopt = kk + 1;
    }
  }
  *hnext = h / scale;
  if (T->kopt >= k && T->kopt != T->kmax && !reduct) {
    fact = dmax(scale / T->alf[T->kopt - 1][T->kopt], SCALMX);
    if (T->acoef[T->kopt + 1] * fact <= wrkmin) {
      *hnext = h / fact;
      T->kopt++;
    }
  }
  return 1;
}

/* driver over one output interval (GKTrajFast.cc:491-538) */
static int ot_odeint(o_traj_t *T, double *ystart, double x1, double x2, double eps,
                     double h1, double hmin) {
  const double TINY = 1.0e-30;
  double y[4], dydx[4], yscal[4];
  double x = x1, hnext = 0.0, hdid;
  double h = (x2 - x1 >= 0.0) ? fabs(h1) : -fabs(h1);
  for (int i = 1; i <= 3; i++) y[i] = ystart[i];
  T->force_reset = 1;
  for (long nstp = 1; nstp <= 10000000L; nstp++) {
    ot_derivs(T, y, dydx);
    if (T->err) return -3;
    for (int i = 1; i <= 3; i++) yscal[i] = fabs(y[i]) + fabs(dydx[i] * h) + TINY;
    if ((x + h - x2) * (x + h - x1) > 0.0) h = x2 - x;
    const int v = ot_bsstep(T, y, dydx, &x, h, eps, yscal, &hdid, &hnext);
    if (T->err) return -3;
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

static int ot_takeastep(o_traj_t *T, double *x, double t1) {
  const double EPS = 1.e-15 * 1000.;
  for (int guard = 0; guard < 100000; guard++) {
    const int v = ot_odeint(T, x, t1, t1 + T->dt, EPS, T->dt, EPS);
    if (v == -3) return -1;
    if (v == 1) return 0;
    if (v == 0) T->dt /= 10.;
    else if (v == -1) { T->dt = 0.0; return 0; }
  }
  o_set_error("trajectory step did not converge");
  return -1;
}

int o_plso_table(double cosi, double a, double *out) {
  o_traj_t T;
  if (ot_init(&T, cosi, a)) return -1;
  out[0] = T.p_lso[1][1]; out[1] = T.p_lso[1][2]; out[2] = T.p_lso[2][1]; out[3] = T.p_lso[2][2];
  return 0;
}

int o_eccentric(double p, double e, double cosi, double a, double dtstart,
                int maxsteps, o_trajdata *traj, int *steps) {
  o_traj_t T;
  *steps = 0;
  if (ot_init(&T, cosi, a)) return -100;
  T.dt = dtstart;
  double ecc_ref = e;
  double x[4] = {0, p, e, cosi};
  double dxdt_est[4];
  T.t = 0;
  int returnflag = 0, done = 0;
  if (p < ot_plso(&T, cosi, e)) { done = 1; returnflag = -1; }
  if (e == 0) { done = 1; returnflag = 1; }
  int i = 0;
  while (!done) {
    o_iekg_t g;
    if (o_iekg_init(&g, x[1], x[2], x[3], a, 1)) return -100;
    const double p3 = g.r3 * (1 - x[2]);
    const double p4 = g.r4 * (1 + x[2]);
    const double alpha_w = g.Q + g.Lz * g.Lz;
    const double beta_w = a * a * (1. - g.E * g.E);
    const double tmp = alpha_w + beta_w;
    double zedminus, betazedplus;
    if (beta_w > 1.e-15) {
      betazedplus = (tmp + sqrt(tmp * tmp - 4. * g.Q * beta_w)) / 2.;
      zedminus = (tmp - sqrt(tmp * tmp - 4. * g.Q * beta_w));
      zedminus /= 2. * beta_w;
      if (fabs(zedminus) < 1.e-15) zedminus = 0.0;
    } else {
      betazedplus = tmp - (g.Q * beta_w / tmp) * (1. - 2. * beta_w / (tmp * tmp));
      zedminus = (g.Q / tmp) * (1. - 2. * beta_w / (tmp * tmp));
    }
    i++;
    o_trajdata *r = &traj[i];
    r->t = T.t; r->p = x[1]; r->ecc = x[2]; r->cosiota = x[3];
    r->E = g.E; r->Lz = g.Lz; r->Q = g.Q; r->p3 = p3; r->p4 = p4;
    r->beta_wilkins = beta_w; r->zedminus = zedminus; r->betazedplus = betazedplus;
    if (i == maxsteps) { done = 1; returnflag = 2; }
    const double p_prev = x[1], ecc_prev = x[2];
    if (ot_takeastep(&T, x, T.t)) return -100;
    if (T.dt * (x[1] - p_prev) > 0.) { done = 1; returnflag = -2; }
    if (x[2] > 0.98 && T.dt * (x[2] - ecc_prev) > 0.) { done = 1; returnflag = -2; }
    ot_derivs(&T, x, dxdt_est);
    if (T.err) return -100;
    while (x[1] + T.dt * dxdt_est[1] < ot_plso(&T, x[3], x[2]) && fabs(T.dt) > 1.e-9) T.dt /= 2.;
    while (x[2] + T.dt * dxdt_est[2] < 0 && fabs(T.dt) > 1.e-9) T.dt /= 2.;
    if (x[2] < 0.5 * ecc_ref) {
      ecc_ref = x[2];
      T.dt /= 2.;
    }
    if (fabs(T.dt) < 1.e-8 || x[2] < 1.e-6) {
      done = 1;
      if (x[2] < 1.e-3 && g.Stable == 1) returnflag = 1;
      else returnflag = 0;
    }
    T.t += T.dt;
  }
  *steps = i;
  return returnflag;
}
