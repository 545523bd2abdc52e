// ref_harness.cpp -- C entry points into the reference's own GSL-free
// trajectory code (GKTrajFast, IEKG, CKG, GKR, Numerical-Recipes numerics),
// compiled from the sources under /root/reference by oracle/Makefile into
// oracle/_ref/libaakref.so.  TEST INFRASTRUCTURE ONLY: the tests use it to
// check the C restatement in oracle/ against the reference bit for bit.
#include <cstring>
#include "Globals.h"
#include "GKTrajFast.h"
#include "IEKG.h"
#include "CKG.h"
#include "GKR.h"
#include "NRCKG.h"
#include "NRRoot.h"

Real rc(const Real x, const Real y);

extern "C" {

int ref_eccentric(double p, double e, double cosi, double a, double dtstart,
                  int maxsteps, double *out, int *steps) {
  GKTrajFast traj(cosi, a);
  traj.p = p;
  traj.ecc = e;
  TrajData *td = new TrajData[maxsteps + 1];
  int n = 0;
  int flag = traj.Eccentric(dtstart, td, maxsteps, n);
  for (int i = 1; i <= n; i++) {
    double *o = out + 12 * (i - 1);
    o[0] = td[i].t; o[1] = td[i].p; o[2] = td[i].ecc; o[3] = td[i].cosiota;
    o[4] = td[i].E; o[5] = td[i].Lz; o[6] = td[i].Q; o[7] = td[i].p3; o[8] = td[i].p4;
    o[9] = td[i].beta_wilkins; o[10] = td[i].zedminus; o[11] = td[i].betazedplus;
  }
  *steps = n;
  delete[] td;
  return flag;
}

void ref_plso_table(double cosi, double a, double *out) {
  GKTrajFast traj(cosi, a);
  out[0] = traj.p_lso[1][1]; out[1] = traj.p_lso[1][2];
  out[2] = traj.p_lso[2][1]; out[3] = traj.p_lso[2][2];
}

void ref_iekg(double p, double e, double cosi, double a, double *out) {
  IEKG g(p, e, cosi, a);
  out[0] = g.E; out[1] = g.Lz; out[2] = g.Q; out[3] = g.Stable;
  out[4] = g.zedminus; out[5] = g.zedplus; out[6] = g.betazedplus;
  out[7] = g.r1; out[8] = g.r2; out[9] = g.r3; out[10] = g.r4;
  out[11] = g.p3; out[12] = g.p4;
  double Om[3];
  g.Frequencies(Om);
  out[13] = Om[0]; out[14] = Om[1]; out[15] = Om[2];
}

void ref_param_derivs(double p, double e, double cosi, double a, double *out) {
  GKTrajFast traj(cosi, a);
  double x[4] = {0, p, e, cosi}, d[4];
  traj.Param_derivs(0.0, x, d);
  out[0] = d[1]; out[1] = d[2]; out[2] = d[3];
}

void ref_ckg(double cosi, double r, double a, double *out) {
  CKG c(0, 0, cosi, USING_cosiota, r, a);
  out[0] = c.E; out[1] = c.Lz; out[2] = c.Q; out[3] = c.Lovercosi;
}

double ref_rf(double x, double y, double z) { return rf(x, y, z); }
double ref_rd(double x, double y, double z) { return rd(x, y, z); }
double ref_rj(double x, double y, double z, double p) { return rj(x, y, z, p); }
double ref_rc(double x, double y) { return rc(x, y); }
double ref_ellf(double phi, double k) { return ellf(phi, k); }
double ref_elle(double phi, double k) { return elle(phi, k); }
double ref_ellpi(double phi, double n, double k) { return ellpi(phi, n, k); }

void ref_quartic_roots(const double *coef, double *re, double *im) {
  Complex a[5], r[5];
  for (int i = 0; i < 5; i++) a[i] = Complex(coef[i], 0.0);
  zroots(a, 4, r, 1, 1.e-15);
  for (int i = 0; i < 4; i++) { re[i] = r[i + 1].real(); im[i] = r[i + 1].imag(); }
}

}  // extern "C"
