// aak_freqmap.hpp -- map Kerr fundamental frequencies to AK parameters.
//
// Solves Omega_Kerr * M_map - Omega_AK(v_map, e, cos iota, s_map) * M = 0 for
// (v_map, M_map, s_map) (KSParMap.cc:323-358,397-440).  The reference calls
// GSL's scaled hybrid solver and stops at sum|f| < 1e-6; this is Powell's
// hybrid method (MINPACK-1 user guide, More, Garbow & Hillstrom 1980):
// dogleg steps in a trust region scaled by the Jacobian column norms,
// forward-difference Jacobian, Broyden rank-one updates (realised for n = 3 by
// refactorising the updated triangle) and a Jacobian refresh after two failed
// steps, with the same stopping rule.
#pragma once

#include "aak_math.hpp"

namespace aak {

struct FreqMap {
  const double *Om;
  double M, e, iota;

  AAK_HD void residual(const double *x, double *f) const {
    const double v = x[0], Mm = x[1], sm = x[2];
    double Or, Ot, Op;
    if (cos(iota) > 0) {
      const double Y = cos(iota);
      const double T = PN::dtdm(v, e, Y, sm);
      Or = PN::drdm(v, e, Y, sm) / T / 2. / kPi;
      Ot = PN::dthetadm(v, e, Y, sm) / T / 2. / kPi;
      Op = PN::dphidm(v, e, Y, sm) / T / 2. / kPi;
    } else {
      const double Y = -cos(iota);
      const double T = PN::dtdm(v, e, Y, -sm);
      Or = PN::drdm(v, e, Y, -sm) / T / 2. / kPi;
      Ot = PN::dthetadm(v, e, Y, -sm) / T / 2. / kPi;
      Op = -PN::dphidm(v, e, Y, -sm) / T / 2. / kPi;
    }
    f[0] = Om[0] * Mm - Or * M;
    f[1] = Om[1] * Mm - Ot * M;
    f[2] = Om[2] * Mm - Op * M;
  }

  AAK_HD static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

  // A = Q R by Householder reflections; cn = column norms of A (may be null).
  AAK_HD static void qr(const double A[3][3], double Q[3][3], double R[3][3], double *cn) {
    double W[3][3];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) W[i][j] = A[i][j];
    if (cn)
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int i = 0; i < 3; i++) s += A[i][j] * A[i][j];
        cn[j] = sqrt(s);
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

  AAK_HD void jacobian(double *x, const double *f, double J[3][3]) const {
    const double eps = 1.4901161193847656e-08;
    for (int j = 0; j < 3; j++) {
      const double xj = x[j];
      double h = eps * fabs(xj);
      if (h == 0.0) h = eps;
      x[j] = xj + h;
      double f1[3];
      residual(x, f1);
      x[j] = xj;
      for (int i = 0; i < 3; i++) J[i][j] = (f1[i] - f[i]) / h;
    }
  }

  AAK_HD static void dogleg(const double R[3][3], const double *diag, const double *qtb, double delta, double *x) {
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
    double g[3] = {0, 0, 0}, w[3];
    for (int j = 0; j < 3; j++) w[j] = diag[j] * x[j];
    const double qnorm = norm3(w);
    if (qnorm <= delta) return;
    for (int j = 0; j < 3; j++) {
      const double t = qtb[j];
      for (int i = j; i < 3; i++) g[i] += R[j][i] * t;
      g[j] /= diag[j];
    }
    const double gnorm = norm3(g);
    double sgnorm = 0.0, alpha = delta / qnorm;
    if (gnorm != 0.0) {
      for (int j = 0; j < 3; j++) g[j] = (g[j] / gnorm) / diag[j];
      for (int j = 0; j < 3; j++) {
        double s = 0.0;
        for (int i = j; i < 3; i++) s += R[j][i] * g[i];
        w[j] = s;
      }
      const double t = norm3(w);
      sgnorm = (gnorm / t) / t;
      alpha = 0.0;
      if (sgnorm < delta) {
        const double bnorm = norm3(qtb);
        double tt = (bnorm / gnorm) * (bnorm / qnorm) * (sgnorm / delta);
        const double dq = delta / qnorm, sd = sgnorm / delta;
        tt = tt - dq * sd * sd + sqrt((tt - dq) * (tt - dq) + (1.0 - dq * dq) * (1.0 - sd * sd));
        alpha = (dq * (1.0 - sd * sd)) / tt;
      }
    }
    const double temp = (1.0 - alpha) * fmin(sgnorm, delta);
    for (int j = 0; j < 3; j++) x[j] = temp * g[j] + alpha * x[j];
  }

  // Hybrid iteration from x (in/out).  Returns iterations used (> 0) or a
  // negative stall code, like the reference's loop (which keeps x either way).
  AAK_HD int solve(double *x, double tol_resid = 1.e-6, int max_iter = 1000) const {
    double f[3], J[3][3], Q[3][3], R[3][3], diag[3], cn[3], qtf[3];
    residual(x, f);
    double fnorm = norm3(f);
    jacobian(x, f, J);
    qr(J, Q, R, cn);
    for (int j = 0; j < 3; j++) diag[j] = (cn[j] == 0.0) ? 1.0 : cn[j];
    const double dx0[3] = {diag[0] * x[0], diag[1] * x[1], diag[2] * x[2]};
    double delta = 100.0 * norm3(dx0);
    if (delta == 0.0) delta = 100.0;
    for (int j = 0; j < 3; j++) qtf[j] = Q[0][j] * f[0] + Q[1][j] * f[1] + Q[2][j] * f[2];
    int iter = 1, ncsuc = 0, ncfail = 0, nslow1 = 0, nslow2 = 0, jeval = 1, it = 0, status = 0;
    for (;;) {
      it++;
      double pd[3], dx[3], xt[3], ft[3], w[3];
      dogleg(R, diag, qtf, delta, pd);
      for (int j = 0; j < 3; j++) {
        dx[j] = -pd[j];
        xt[j] = x[j] + dx[j];
      }
      const double dd[3] = {diag[0] * dx[0], diag[1] * dx[1], diag[2] * dx[2]};
      const double pnorm = norm3(dd);
      if (iter == 1) delta = fmin(delta, pnorm);
      residual(xt, ft);
      if (!(is_finite(ft[0]) && is_finite(ft[1]) && is_finite(ft[2]))) {
        status = -2;
        break;
      }
      const double fnorm1 = norm3(ft);
      double actred = -1.0;
      if (fnorm1 < fnorm) actred = 1.0 - (fnorm1 / fnorm) * (fnorm1 / fnorm);
      for (int i = 0; i < 3; i++) {
        double s = 0.0;
        for (int j = i; j < 3; j++) s += R[i][j] * dx[j];
        w[i] = qtf[i] + s;
      }
      const double tn = norm3(w);
      double prered = 0.0;
      if (tn < fnorm) prered = 1.0 - (tn / fnorm) * (tn / fnorm);
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
      const bool accepted = (ratio >= 1e-4);
      if (accepted) {
        for (int j = 0; j < 3; j++) {
          x[j] = xt[j];
          f[j] = ft[j];
        }
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
          jacobian(x, f, J);
          qr(J, Q, R, cn);
          for (int j = 0; j < 3; j++) diag[j] = fmax(diag[j], cn[j]);
          for (int j = 0; j < 3; j++) qtf[j] = Q[0][j] * f[0] + Q[1][j] * f[1] + Q[2][j] * f[2];
          jeval = 1;
        } else {
          double u[3], vv[3], qtft[3];
          for (int j = 0; j < 3; j++) {
            qtft[j] = Q[0][j] * ft[0] + Q[1][j] * ft[1] + Q[2][j] * ft[2];
            u[j] = (qtft[j] - w[j]) / pnorm;
            vv[j] = diag[j] * ((diag[j] * dx[j]) / pnorm);
            if (accepted) qtf[j] = qtft[j];
          }
          double Mt[3][3], Q2[3][3], R2[3][3], Qn[3][3], qn[3];
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Mt[i][j] = R[i][j] + u[i] * vv[j];
          qr(Mt, Q2, R2, nullptr);
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) Qn[i][j] = Q[i][0] * Q2[0][j] + Q[i][1] * Q2[1][j] + Q[i][2] * Q2[2][j];
          for (int j = 0; j < 3; j++) qn[j] = Q2[0][j] * qtf[0] + Q2[1][j] * qtf[1] + Q2[2][j] * qtf[2];
          for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
              Q[i][j] = Qn[i][j];
              R[i][j] = R2[i][j];
            }
          for (int j = 0; j < 3; j++) qtf[j] = qn[j];
          jeval = 0;
        }
      }
      if (status) break;
      if (fabs(f[0]) + fabs(f[1]) + fabs(f[2]) < tol_resid) break;
      if (it >= max_iter) break;
    }
    return status ? status : it;
  }

  // Whether a polished root stayed near its learned start (each component
  // within `rel` of it, relative): a Newton run that wandered further may have
  // found another root of the system and is not trusted.
  AAK_HD static bool near_start(const double *x, const double *x0, double rel) {
    for (int j = 0; j < 3; j++)
      if (!(fabs(x[j] - x0[j]) <= rel * fmax(fabs(x0[j]), 1e-3))) return false;
    return true;
  }

  // Newton iteration from a close start x (in/out): forward-difference
  // Jacobian (as the hybrid's), each step solved by QR.  Stops once
  // sum|f| <= tol; returns the Newton steps taken (0 if x already met tol) or
  // -1 when max_iter steps do not meet it or a step leaves the finite range
  // (x is then left at its last finite iterate).
  AAK_HD int polish(double *x, double tol, int max_iter) const {
    double f[3];
    residual(x, f);
    for (int it = 0;; it++) {
      if (!(is_finite(f[0]) && is_finite(f[1]) && is_finite(f[2]))) return -1;
      if (fabs(f[0]) + fabs(f[1]) + fabs(f[2]) <= tol) return it;
      if (it >= max_iter) return -1;
      double J[3][3], Q[3][3], R[3][3], qtf[3], dx[3];
      jacobian(x, f, J);
      qr(J, Q, R, nullptr);
      for (int j = 0; j < 3; j++) qtf[j] = Q[0][j] * f[0] + Q[1][j] * f[1] + Q[2][j] * f[2];
      for (int j = 2; j >= 0; j--) {
        double s = qtf[j];
        for (int k = j + 1; k < 3; k++) s -= R[j][k] * dx[k];
        if (R[j][j] == 0.0) return -1;
        dx[j] = s / R[j][j];
      }
      double xt[3], ft[3];
      for (int j = 0; j < 3; j++) xt[j] = x[j] - dx[j];
      residual(xt, ft);
      if (!(is_finite(ft[0]) && is_finite(ft[1]) && is_finite(ft[2]))) return -1;
      for (int j = 0; j < 3; j++) {
        x[j] = xt[j];
        f[j] = ft[j];
      }
    }
  }
};

}  // namespace aak
