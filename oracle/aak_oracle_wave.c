/*
 * aak_oracle_wave.c -- AAK driver of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see aak_oracle.h).  Restates AAKpy.cc:16-51
 * (sparse stage), AAK.cc:27-318 (PNevolution: AK evolution, fit, plunge,
 * spline onto the dense grid, sequential Adams-Bashforth phase folds) and
 * AAK.cc:321-591 (harmonic mode sum, LISA response, Planck taper, GenWave).
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "aak_oracle_int.h"

static int imax2(int a, int b) { return a > b ? a : b; }
static int imin2(int a, int b) { return a < b ? a : b; }

int o_sparse_stage(double T_fit, int length, double dt, double iota, double s,
                   double p, double e, double M, double mu, double psi,
                   double *e_traj, double *v_map, double *M_map, double *s_map,
                   double *dt_map, double *Phi0, int *ecc_flag) {
  const int maxsteps = 100;
  o_trajdata traj[101];
  int steps = 0;
  const double tr = length * dt / O_MSUN_SEC / M / M * mu;
  const double dt_fit = ((tr < T_fit) ? tr : T_fit) / (maxsteps - 1);
  const int flag = o_eccentric(p, e, cos(iota), s, dt_fit, maxsteps, traj, &steps);
  if (ecc_flag) *ecc_flag = flag;
  if (flag == -100) return -1;
  if (steps < 3) { o_set_error("trajectory too short (p0 at or inside the LSO?)"); return -2; }
  for (int i = 1; i <= steps; i++) {
    o_iekg_t g;
    if (o_iekg_init(&g, traj[i].p, traj[i].ecc, traj[i].cosiota, s, 1)) return -1;
    double Om[3], map[3];
    o_iekg_frequencies(&g, Om);
    if (i == 1) *Phi0 = o_parang_phi0(e, psi);
    const int st = o_parmap(map, Om, traj[i].p, M, s, traj[i].ecc, iota);
    (void)st;
    e_traj[i - 1] = traj[i].ecc;
    v_map[i - 1] = map[0];
    M_map[i - 1] = map[1];
    s_map[i - 1] = map[2];
    dt_map[i - 1] = traj[i].t * O_MSUN_SEC * M * M / mu;
  }
  return steps;
}

int o_pn_j0(double timestep, int vlength, const double *dt_map, int steps) {
  const double dt_large = dt_map[1];
  return (int)((vlength * timestep - dt_map[steps - 1]) / dt_large + steps + 1);
}

static double o_pn_p(double v, double e, double coslam, double S, double M, double Mphys) {
  return (1. - e * e) * pow(o_omegaphi(v, e, coslam, S, M) * Mphys * O_MSUN_SEC, -2. / 3.);
}

int o_pn_knots(double timestep, int vlength, const double *par,
               const double *e_traj, const double *v_map, double M_phys,
               const double *M_map, double S_phys, const double *S_map,
               const double *dt_map, int steps, int backint,
               double *t_in, double *e_in, double *v_in, double *M_in,
               double *S_in, int cap, int *info) {
  const double mu = par[0], M0 = par[1], S0 = par[2], e0 = par[3];
  const double v0 = v_map[0];
  const double coslam = cos(par[4]);
  double edot = 0, vdot = 0, edotm, vdotm, edotp, vdotp;
  const double dt_large = dt_map[1];
  const int j0 = o_pn_j0(timestep, vlength, dt_map, steps);
  int j_start = j0, j_min = j0, j_end = 2 * j0 - 1, j_max = j_end, j_max_temp = j_max;
  double t_end = vlength * timestep;
  if (j0 < steps || j0 > 400000000) { o_set_error("knot count out of range"); return -3; }
  double *e_AK = calloc((size_t)2 * j0, sizeof(double));
  double *v_AK = calloc((size_t)2 * j0, sizeof(double));
  double *t_fit = calloc((size_t)2 * j0, sizeof(double));
  double *e_fit = calloc((size_t)2 * j0, sizeof(double));
  double *v_fit = calloc((size_t)2 * j0, sizeof(double));
  double *M_fit = calloc((size_t)2 * j0, sizeof(double));
  double *S_fit = calloc((size_t)2 * j0, sizeof(double));
  int rc = 0;
  e_AK[j0] = e0;
  v_AK[j0] = v0;
  /* AK evolution over the sparse-trajectory times (AAK.cc:65-81) */
  for (int j = j0; j < j0 + steps - 1; j++) {
    edotm = edot;
    vdotm = vdot;
    edot = o_dedt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
    vdot = o_dvdt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
    if (j == j0) { edotm = edot; vdotm = vdot; }
    e_AK[j + 1] = e_AK[j] + (1.5 * edot - .5 * edotm) * (dt_map[j - j0 + 1] - dt_map[j - j0]);
    v_AK[j + 1] = v_AK[j] + (1.5 * vdot - .5 * vdotm) * (dt_map[j - j0 + 1] - dt_map[j - j0]);
    if (e_AK[j + 1] < 0. || v_AK[j + 1] < v_AK[j] || isnan(e_AK[j + 1]) || isnan(v_AK[j + 1])) {
      if (j_max == j_max_temp) j_max = j;
      e_AK[j + 1] = e_AK[j_max];
      v_AK[j + 1] = v_AK[j_max];
    }
  }
  j_max_temp = j_max;
  /* fit coefficients (AAK.cc:86-106) */
  int points = imin2(steps, j_max - j0 + 1);
  while (points >= 2 && e_traj[points - 1] > e_traj[points - 2]) points--;
  if (points < 3) { o_set_error("Fitting error: t_0 too close to plunge"); rc = -4; goto done; }
  {
    double e_coeff[2], v_coeff[2], M_coeff[2], S_coeff[2];
    double *e_diff = malloc(sizeof(double) * points), *v_diff = malloc(sizeof(double) * points);
    for (int i = 0; i < points; i++) {
      e_diff[i] = e_traj[i] - e_AK[i + j0];
      v_diff[i] = v_map[i] - v_AK[i + j0];
    }
    o_polyfit(e_coeff, dt_map, e_diff, points);
    o_polyfit(v_coeff, dt_map, v_diff, points);
    o_polyfit(M_coeff, dt_map, M_map, points);
    o_polyfit(S_coeff, dt_map, S_map, points);
    free(e_diff);
    free(v_diff);
    /* AK evolution on the uniform grid (AAK.cc:110-122) */
    for (int j = j0 + steps - 1; j < j_end; j++) {
      edotm = edot;
      vdotm = vdot;
      edot = o_dedt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
      vdot = o_dvdt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
      e_AK[j + 1] = e_AK[j] + (1.5 * edot - .5 * edotm) * dt_large;
      v_AK[j + 1] = v_AK[j] + (1.5 * vdot - .5 * vdotm) * dt_large;
      if (e_AK[j + 1] < 0. || v_AK[j + 1] < v_AK[j] || isnan(e_AK[j + 1]) || isnan(v_AK[j + 1])) {
        if (j_max == j_max_temp) j_max = j;
        e_AK[j + 1] = e_AK[j_max];
        v_AK[j + 1] = v_AK[j_max];
      }
    }
    j_max_temp = j_max;
    /* fitted knots (AAK.cc:127-144) */
    for (int j = j0; j <= j_end; j++) {
      double dt;
      if (j < j0 + steps) dt = dt_map[j - j0];
      else dt = dt_map[steps - 1] + dt_large * (j - j0 - steps + 1);
      const double dt2 = dt * dt;
      t_fit[j] = dt;
      e_fit[j] = e_AK[j] + e_coeff[0] * dt + e_coeff[1] * dt2;
      v_fit[j] = v_AK[j] + v_coeff[0] * dt + v_coeff[1] * dt2;
      M_fit[j] = M0 + (M_coeff[0] * dt + M_coeff[1] * dt2) * O_MSUN_SEC;
      S_fit[j] = S0 + S_coeff[0] * dt + S_coeff[1] * dt2;
      if (e_fit[j] < 0. || v_fit[j] < v_fit[imax2(j0, j - 1)]) {
        if (j_max == j_max_temp) j_max = j - 1;
        e_fit[j] = e_fit[j_max];
        v_fit[j] = v_fit[j_max];
        M_fit[j] = M_fit[j_max];
        S_fit[j] = S_fit[j_max];
      }
    }
    j_max_temp = j_max;
    /* plunge search (AAK.cc:149-173) */
    const int j_RR = (int)((O_MSUN_SEC * M_phys * O_MSUN_SEC * M_phys / mu) / dt_large) + 1;
    const double z1 = 1. + pow(1. - S_phys * S_phys, 1. / 3.) * (pow(1. + S_phys, 1. / 3.) + pow(1. - S_phys, 1. / 3.));
    const double z2 = sqrt(3. * S_phys * S_phys + z1 * z1);
    const double LSO_min = 3. + z2 - sqrt((3. - z1) * (3. + z1 + 2. * z2));
    const double LSO_max = 3. + z2 + sqrt((3. - z1) * (3. + z1 + 2. * z2));
    for (int j = j0; j <= j_max_temp; j++) {
      if (j > j_max) break;
      if (o_pn_p(v_fit[j], e_fit[j], coslam, S_fit[j], M_fit[j], M_phys) < LSO_min && j_max == j_max_temp) {
        j_min = imax2(j - j_RR, j0);
        j_max = imin2(j + 1, j_max_temp);
      }
      if (j > j0 && (j - j0) % j_RR == 0 &&
          o_pn_p(v_fit[j], e_fit[j], coslam, S_fit[j], M_fit[j], M_phys) < LSO_max && j_max == j_max_temp) {
        j_min = imax2(j - j_RR, j0);
        o_iekg_t g;
        if (o_iekg_init(&g, o_pn_p(v_fit[j], e_fit[j], coslam, S_fit[j], M_fit[j], M_phys), e_fit[j], coslam, S_phys, 1)) { rc = -5; goto done; }
        if (g.Stable == -1 || g.E > 1) j_max = imin2(j + 1, j_max_temp);
      }
    }
    while (j_max - j_min > 1) {
      const int j_mid = (j_max + j_min) / 2;
      o_iekg_t g;
      if (o_iekg_init(&g, o_pn_p(v_fit[j_mid], e_fit[j_mid], coslam, S_fit[j_mid], M_fit[j_mid], M_phys), e_fit[j_mid], coslam, S_phys, 1)) { rc = -5; goto done; }
      if (g.Stable == -1 || g.E > 1) j_max = j_mid;
      else j_min = j_mid;
    }
    /* backward extension (AAK.cc:177-209) */
    if (backint) {
      j_end = j_max;
      t_end = (t_fit[j_end] < vlength * timestep) ? t_fit[j_end] : vlength * timestep;
      j_start = j0 + floor((t_end - vlength * timestep) / dt_large);
      for (int j = j0; j > j_start; j--) {
        edotp = edot;
        vdotp = vdot;
        edot = o_dedt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
        vdot = o_dvdt(v_AK[j], e_AK[j], coslam, mu, M0, S0);
        if (j == j0) { edotp = edot; vdotp = vdot; }
        e_AK[j - 1] = e_AK[j] - (1.5 * edot - .5 * edotp) * dt_large;
        v_AK[j - 1] = v_AK[j] - (1.5 * vdot - .5 * vdotp) * dt_large;
      }
      for (int j = j_start; j < j0; j++) {
        const double dt_und = dt_large * (j - j0);
        double dt;
        if (j > j0 - points) dt = (dt_und + dt_und * (j - j0 + points) / points) / 2.;
        else dt = -dt_large * points / 2.;
        const double dt2 = dt * dt;
        t_fit[j] = dt_und;
        e_fit[j] = e_AK[j] + e_coeff[0] * dt + e_coeff[1] * dt2;
        v_fit[j] = v_AK[j] + v_coeff[0] * dt + v_coeff[1] * dt2;
        M_fit[j] = M0 + (M_coeff[0] * dt + M_coeff[1] * dt2) * O_MSUN_SEC;
        S_fit[j] = S0 + S_coeff[0] * dt + S_coeff[1] * dt2;
      }
    }
  }
  {
    const int K = j_end - j_start + 1;
    if (K > cap) { o_set_error("knot buffer too small"); rc = -6; goto done; }
    for (int j = j_start; j <= j_end; j++) {
      t_in[j - j_start] = t_fit[j];
      e_in[j - j_start] = e_fit[j];
      v_in[j - j_start] = v_fit[j];
      M_in[j - j_start] = M_fit[j];
      S_in[j - j_start] = S_fit[j];
    }
    info[0] = vlength - (int)(t_end / timestep);
    info[1] = j0;
    info[2] = j_min;
    info[3] = j_max;
    info[4] = j_max_temp;
    info[5] = j_start;
    info[6] = j_end;
    info[7] = (j_max != j_max_temp);
    info[8] = points;
    rc = K;
  }
done:
  free(e_AK); free(v_AK); free(t_fit); free(e_fit); free(v_fit); free(M_fit); free(S_fit);
  return rc;
}

/* Dense arrays and sequential phase folds (AAK.cc:228-295).
 * dense layout: 10 blocks of (vlength+1): t,e,v,M,S,gim,Phi,alp,nu,gimdot. */
static int o_pn_dense_rates(double timestep, int vlength, const double *par, int backint,
                            const double *t_in, const double *e_in, const double *v_in,
                            const double *M_in, const double *S_in, int K, int i0,
                            int plunged, double t_fit_jmin, double *dense,
                            int *i_plunge, int *i_buffer, double *Phidotvec, double *alpdotvec);

static int o_pn_dense(double timestep, int vlength, const double *par, int backint,
                      const double *t_in, const double *e_in, const double *v_in,
                      const double *M_in, const double *S_in, int K, int i0,
                      int plunged, double t_fit_jmin, double *dense,
                      int *i_plunge, int *i_buffer) {
  return o_pn_dense_rates(timestep, vlength, par, backint, t_in, e_in, v_in, M_in, S_in, K, i0, plunged,
                          t_fit_jmin, dense, i_plunge, i_buffer, NULL, NULL);
}

/* Same fold; Phidotvec/alpdotvec (optional, vlength+1) receive the rates as
 * PNevolutionPhase stores them (AAKPhase.cc:233-290). */
static int o_pn_dense_rates(double timestep, int vlength, const double *par, int backint,
                            const double *t_in, const double *e_in, const double *v_in,
                            const double *M_in, const double *S_in, int K, int i0,
                            int plunged, double t_fit_jmin, double *dense,
                            int *i_plunge, int *i_buffer, double *Phidotvec, double *alpdotvec) {
  const size_t L = (size_t)vlength + 1;
  double *t = dense, *e = dense + L, *v = dense + 2 * L, *M = dense + 3 * L, *S = dense + 4 * L;
  double *gim = dense + 5 * L, *Phi = dense + 6 * L, *alp = dense + 7 * L, *nu = dense + 8 * L, *gdv = dense + 9 * L;
  const double coslam = cos(par[4]);
  (void)backint;
  t[0] = t_in[0];
  for (int i = 1; i < vlength; i++) t[i] = (i - i0) * timestep;
  t[vlength] = t_in[K - 1];
  const double *ys[4] = {e_in, v_in, M_in, S_in};
  double *outs[4] = {e, v, M, S};
  double *c = malloc(sizeof(double) * K);
  for (int a = 0; a < 4; a++) {
    if (o_spline_natural(t_in, ys[a], K, c)) { free(c); return -1; }
    for (size_t i = 0; i < L; i++) outs[a][i] = o_spline_eval(t_in, ys[a], c, K, t[i]);
  }
  free(c);
  if (!plunged) *i_plunge = vlength - 1;
  else *i_plunge = i0 + (int)(t_fit_jmin / timestep);
  if (*i_plunge < 0 || *i_plunge > vlength) { o_set_error("plunge index outside the waveform"); return -2; }
  const int ip = *i_plunge;
  *i_buffer = (int)(10. / (o_omegaphi(v[ip], e[ip], coslam, S[ip], M[ip]) / 2. / M_PI) / timestep) + 1;
  const int ib = *i_buffer;
  double gimdot = 0, Phidot = 0, alpdot = 0, gimdotm, Phidotm, alpdotm, gimdotp, Phidotp, alpdotp;
  gim[i0] = par[5];
  Phi[i0] = par[6];
  alp[i0] = par[11];
  for (int i = i0; i < vlength; i++) {
    gimdotm = gimdot;
    Phidotm = Phidot;
    alpdotm = alpdot;
    gimdot = (o_dthetadm(v[i], e[i], coslam, S[i]) - o_drdm(v[i], e[i], coslam, S[i])) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    Phidot = o_drdm(v[i], e[i], coslam, S[i]) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    alpdot = (o_dphidm(v[i], e[i], coslam, S[i]) - o_dthetadm(v[i], e[i], coslam, S[i])) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    if (Phidot <= 0.) {
      if (i == i0) { o_set_error("non-positive radial frequency at t0"); return -3; }
      Phidot = Phidotm;
    }
    if (i == i0) { gimdotm = gimdot; Phidotm = Phidot; alpdotm = alpdot; }
    nu[i] = Phidot / 2. / M_PI;
    gdv[i] = gimdot;
    if (Phidotvec) { Phidotvec[i] = Phidot; alpdotvec[i] = alpdot; }
    gim[i + 1] = gim[i] + (1.5 * gimdot - .5 * gimdotm) * timestep;
    Phi[i + 1] = Phi[i] + (1.5 * Phidot - .5 * Phidotm) * timestep;
    alp[i + 1] = alp[i] + (1.5 * alpdot - .5 * alpdotm) * timestep;
    if (i > ip + ib) {
      gim[i + 1] = gim[ip];
      Phi[i + 1] = Phi[ip];
      alp[i + 1] = alp[ip];
    }
  }
  {
    const int n = vlength;
    nu[n] = o_drdm(v[n], e[n], coslam, S[n]) / o_dtdm(v[n], e[n], coslam, S[n]) / (2. * M_PI * M[n]);
    gdv[n] = (o_dthetadm(v[n], e[n], coslam, S[n]) - o_drdm(v[n], e[n], coslam, S[n])) / o_dtdm(v[n], e[n], coslam, S[n]) / M[n];
    if (Phidotvec) {
      Phidotvec[n] = o_drdm(v[n], e[n], coslam, S[n]) / o_dtdm(v[n], e[n], coslam, S[n]) / M[n];
      alpdotvec[n] = (o_dphidm(v[n], e[n], coslam, S[n]) - o_dthetadm(v[n], e[n], coslam, S[n])) / o_dtdm(v[n], e[n], coslam, S[n]) / M[n];
    }
  }
  for (int i = i0; i > 0; i--) {
    gimdotp = gimdot;
    Phidotp = Phidot;
    alpdotp = alpdot;
    gimdot = (o_dthetadm(v[i], e[i], coslam, S[i]) - o_drdm(v[i], e[i], coslam, S[i])) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    Phidot = o_drdm(v[i], e[i], coslam, S[i]) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    alpdot = (o_dphidm(v[i], e[i], coslam, S[i]) - o_dthetadm(v[i], e[i], coslam, S[i])) / o_dtdm(v[i], e[i], coslam, S[i]) / M[i];
    if (i == i0) { gimdotp = gimdot; Phidotp = Phidot; alpdotp = alpdot; }
    nu[i] = Phidot / 2. / M_PI;
    gdv[i] = gimdot;
    if (Phidotvec) { Phidotvec[i] = Phidot; alpdotvec[i] = alpdot; }
    gim[i - 1] = gim[i] - (1.5 * gimdot - .5 * gimdotp) * timestep;
    Phi[i - 1] = Phi[i] - (1.5 * Phidot - .5 * Phidotp) * timestep;
    alp[i - 1] = alp[i] - (1.5 * alpdot - .5 * alpdotp) * timestep;
  }
  nu[0] = o_drdm(v[0], e[0], coslam, S[0]) / o_dtdm(v[0], e[0], coslam, S[0]) / (2. * M_PI * M[0]);
  gdv[0] = (o_dthetadm(v[0], e[0], coslam, S[0]) - o_drdm(v[0], e[0], coslam, S[0])) / o_dtdm(v[0], e[0], coslam, S[0]) / M[0];
  if (Phidotvec) {
    Phidotvec[0] = o_drdm(v[0], e[0], coslam, S[0]) / o_dtdm(v[0], e[0], coslam, S[0]) / M[0];
    alpdotvec[0] = (o_dphidm(v[0], e[0], coslam, S[0]) - o_dthetadm(v[0], e[0], coslam, S[0])) / o_dtdm(v[0], e[0], coslam, S[0]) / M[0];
  }
  return 0;
}

/* Peters-Mathews harmonic coefficients a_n, b_n, c_n of the AK mode sum
 * (AAK.cc:493-504) at harmonic n, eccentricity e, amplitude Amp and phase nPhi. */
void o_mode_abc(int n, double e, double Amp, double nPhi, double *abc) {
  double J[5], Jb[4];
  const double ne = n * e;
  if (n == 1) {
    o_bessel_jn_array(0, 3, ne, Jb);
    J[0] = -Jb[1];
    J[1] = Jb[0];
    J[2] = Jb[1];
    J[3] = Jb[2];
    J[4] = Jb[3];
  } else {
    o_bessel_jn_array(n - 2, n + 2, ne, J);
  }
  const double cnp = cos(nPhi), snp = sin(nPhi);
  abc[0] = -n * Amp * (J[0] - 2. * e * J[1] + 2. / n * J[2] + 2. * e * J[3] - J[4]) * cnp;
  abc[1] = -n * Amp * sqrt(1 - e * e) * (J[0] - 2. * J[2] + J[4]) * snp;
  abc[2] = 2. * Amp * J[2] * cnp;
}

int o_waveform_from_dense(const double *dense, int vlength, double timestep,
                          int nmodes, double zeta, double M_phys, double lam,
                          double qS, double phiS, double qK, double phiK,
                          int mich, int i_plunge, int i_buffer,
                          double *t, double *hI, double *hII) {
  const size_t L = (size_t)vlength + 1;
  const double *tvec = dense, *evec = dense + L, *vvec = dense + 2 * L, *Mvec = dense + 3 * L, *Svec = dense + 4 * L;
  const double *gimvec = dense + 5 * L, *Phivec = dense + 6 * L, *alpvec = dense + 7 * L, *nuvec = dense + 8 * L, *gdvec = dense + 9 * L;
  const double coslam = cos(lam), sinlam = sin(lam);
  const double cosqS = cos(qS), sinqS = sin(qS), cosqK = cos(qK), sinqK = sin(qK);
  const double cosphiK = cos(phiK), sinphiK = sin(phiK);
  const double halfsqrt3 = sqrt(3.) / 2.;
  double J[5];
  double *Jb = malloc(sizeof(double) * 8);
  for (int i = 0; i < vlength; i++) {
    t[i] = tvec[i];
    hI[i] = 0.;
    hII[i] = 0.;
    const double e = evec[i], v = vvec[i], M = Mvec[i], S = Svec[i];
    const double gim = gimvec[i], Phi = Phivec[i], alp = alpvec[i], nu = nuvec[i], gimdot = gdvec[i];
    const double cosalp = cos(alp), sinalp = sin(alp);
    const double cosqL = cosqK * coslam + sinqK * sinlam * cosalp;
    const double sinqL = sqrt(1. - cosqL * cosqL);
    const double phiLup = sinqK * sinphiK * coslam - cosphiK * sinlam * sinalp - cosqK * sinphiK * sinlam * cosalp;
    const double phiLdown = sinqK * cosphiK * coslam + sinphiK * sinlam * sinalp - cosqK * cosphiK * sinlam * cosalp;
    const double phiL = atan2(phiLup, phiLdown);
    const double Ldotn = cosqL * cosqS + sinqL * sinqS * cos(phiL - phiS);
    const double Ldotn2 = Ldotn * Ldotn;
    const double Sdotn = cosqK * cosqS + sinqK * sinqS * cos(phiK - phiS);
    const double betaup = -Sdotn + coslam * Ldotn;
    const double betadown = sinqS * sin(phiK - phiS) * sinlam * cosalp + (cosqK * Sdotn - cosqS) / sinqK * sinlam * sinalp;
    const double beta = atan2(betaup, betadown);
    const double gam = 2. * (gim + beta);
    const double cos2gam = cos(gam), sin2gam = sin(gam);
    double orbphs = 0, cosorbphs = 0, sinorbphs, FplusI, FcrosI, FplusII, FcrosII;
    if (mich) {
      orbphs = 2. * M_PI * t[i] / O_YEAR;
      cosorbphs = cos(orbphs - phiS);
      sinorbphs = sin(orbphs - phiS);
      const double cosq = .5 * cosqS - halfsqrt3 * sinqS * cosorbphs;
      const double phiw = orbphs + atan2(halfsqrt3 * cosqS + .5 * sinqS * cosorbphs, sinqS * sinorbphs);
      const double psiup = .5 * cosqK - halfsqrt3 * sinqK * cos(orbphs - phiK) - cosq * (cosqK * cosqS + sinqK * sinqS * cos(phiK - phiS));
      const double psidown = .5 * sinqK * sinqS * sin(phiK - phiS) - halfsqrt3 * cos(orbphs) * (cosqK * sinqS * sin(phiS) - cosqS * sinqK * sin(phiK)) -
                             halfsqrt3 * sin(orbphs) * (cosqS * sinqK * cos(phiK) - cosqK * sinqS * cos(phiS));
      const double psi = atan2(psiup, psidown);
      const double cosq1 = .5 * (1. + cosq * cosq);
      const double cos2phi = cos(2. * phiw), sin2phi = sin(2. * phiw);
      const double cos2psi = cos(2. * psi), sin2psi = sin(2. * psi);
      FplusI = cosq1 * cos2phi * cos2psi - cosq * sin2phi * sin2psi;
      FcrosI = cosq1 * cos2phi * sin2psi + cosq * sin2phi * cos2psi;
      FplusII = cosq1 * sin2phi * cos2psi + cosq * cos2phi * sin2psi;
      FcrosII = cosq1 * sin2phi * sin2psi - cosq * cos2phi * cos2psi;
    } else {
      const double up_ldc = (cosqS * sinqK * cos(phiS - phiK) - cosqK * sinqS);
      const double dw_ldc = (sinqK * sin(phiS - phiK));
      double psi_ldc;
      if (dw_ldc != 0.0) psi_ldc = -atan2(up_ldc, dw_ldc);
      else psi_ldc = 0.5 * M_PI;
      const double c2 = cos(2. * psi_ldc), s2 = sin(2. * psi_ldc);
      FplusI = c2;
      FcrosI = -s2;
      FplusII = s2;
      FcrosII = c2;
    }
    const double Amp = pow(o_omegaphi(v, e, coslam, S, M) * M_phys * O_MSUN_SEC, 2. / 3.) * zeta;
    double rot[4];
    o_rotcoeff(rot, lam, qS, phiS, qK, phiK, alp);
    for (int n = 1; n <= nmodes; n++) {
      double nPhi;
      if (mich) {
        const double fn = n * nu + gimdot / M_PI;
        const double Doppler = 2. * M_PI * fn * O_AUSEC * sinqS * cosorbphs;
        nPhi = n * Phi + Doppler;
      } else {
        nPhi = n * Phi;
      }
      double abc[3];
      o_mode_abc(n, e, Amp, nPhi, abc);
      const double a = abc[0], b = abc[1], c = abc[2];
      const double Aplus0 = -(1. + Ldotn2) * (a * cos2gam - b * sin2gam) + c * (1 - Ldotn2);
      const double Acros0 = 2. * Ldotn * (b * cos2gam + a * sin2gam);
      const double Aplus = Aplus0 * rot[0] + Acros0 * rot[1];
      const double Acros = Aplus0 * rot[2] + Acros0 * rot[3];
      double hnI, hnII;
      if (mich) {
        hnI = halfsqrt3 * (FplusI * Aplus + FcrosI * Acros);
        hnII = halfsqrt3 * (FplusII * Aplus + FcrosII * Acros);
      } else {
        hnI = FplusI * Aplus + FcrosI * Acros;
        hnII = FplusII * Aplus + FcrosII * Acros;
      }
      hI[i] += hnI;
      hII[i] += hnII;
    }
    if (i == i_plunge + i_buffer) break;
  }
  free(Jb);
  /* Planck taper (AAK.cc:536-547) */
  const double t_plunge = t[i_plunge], t_zero = t_plunge + timestep * i_buffer;
  for (int i = i_plunge + 1; i < vlength; i++) {
    t[i] = tvec[i];
    if (i < i_plunge + i_buffer) {
      hI[i] = hI[i] / (exp((t_plunge - t_zero) / (t[i] - t_plunge) + (t_plunge - t_zero) / (t[i] - t_zero)) + 1.);
      hII[i] = hII[i] / (exp((t_plunge - t_zero) / (t[i] - t_plunge) + (t_plunge - t_zero) / (t[i] - t_zero)) + 1.);
    } else {
      hI[i] = 0.;
      hII[i] = 0.;
    }
  }
  return 0;
}

int o_aak_wave(const double *pars, double T_fit, int length, double dt,
               int LISA, int backint, int n_max,
               double *t, double *hI, double *hII, double *dense, double *meta) {
  const double iota = pars[0], s = pars[1], p = pars[2], e = pars[3], M = pars[4], mu = pars[5];
  const double gamma = pars[6], psi = pars[7], alph = pars[8], qS = pars[9], phiS = pars[10];
  const double qK = pars[11], phiK = pars[12], D = pars[13];
  double e_traj[100], v_map[100], M_map[100], s_map[100], dt_map[100], Phi0 = 0;
  const int steps = o_sparse_stage(T_fit, length, dt, iota, s, p, e, M, mu, psi,
                                   e_traj, v_map, M_map, s_map, dt_map, &Phi0, NULL);
  if (steps < 3) return steps < 0 ? steps : -2;
  double par[12];
  par[0] = mu * O_MSUN_SEC;
  par[1] = M_map[0] * O_MSUN_SEC;
  par[2] = s_map[0];
  par[3] = e_traj[0];
  par[4] = iota;
  par[5] = gamma;
  par[6] = Phi0;
  par[7] = qS;
  par[8] = phiS;
  par[9] = qK;
  par[10] = phiK;
  par[11] = alph;
  int nmodes = (int)(30 * par[3]);
  if (par[3] < 0.135) nmodes = 4;
  if (n_max > 0) nmodes = n_max;
  const double zeta = par[0] / D / O_GPC;
  const int j0 = o_pn_j0(dt, length, dt_map, steps);
  if (j0 < steps || j0 > 400000000) { o_set_error("knot count out of range"); return -3; }
  const int cap = 2 * j0 + 8;
  double *kn = malloc(sizeof(double) * 5 * (size_t)cap);
  int info[9];
  const int K = o_pn_knots(dt, length, par, e_traj, v_map, M, M_map, s, s_map, dt_map, steps, backint,
                           kn, kn + cap, kn + 2 * cap, kn + 3 * cap, kn + 4 * cap, cap, info);
  if (K < 0) { free(kn); return K; }
  const size_t L = (size_t)length + 1;
  double *dn = dense ? dense : malloc(sizeof(double) * 10 * L);
  /* t_fit[j_min] is the knot time at j_min (index relative to j_start) */
  const double t_fit_jmin = kn[info[2] - info[5]];
  int i_plunge = 0, i_buffer = 0;
  int rc = o_pn_dense(dt, length, par, backint, kn, kn + cap, kn + 2 * cap, kn + 3 * cap, kn + 4 * cap,
                      K, info[0], info[7], t_fit_jmin, dn, &i_plunge, &i_buffer);
  if (rc == 0)
    rc = o_waveform_from_dense(dn, length, dt, nmodes, zeta, M, iota, qS, phiS, qK, phiK, LISA,
                               i_plunge, i_buffer, t, hI, hII);
  if (meta) {
    meta[0] = steps; meta[1] = nmodes; meta[2] = i_plunge; meta[3] = i_buffer;
    meta[4] = K; meta[5] = info[0]; meta[6] = Phi0; meta[7] = zeta;
  }
  if (!dense) free(dn);
  free(kn);
  return rc;
}

/* AAK phases and frequencies (AAKpy.cc:53-88, AAKPhase.cc:13-374): the sparse
 * stage and knots of the waveform, the same dense folds, then per sample
 * t, Phi, Phi+gamma, Phi+gamma+alpha, their rates and e (GenPhase,
 * AAKPhase.cc:338-346).  out: 8 arrays of length each. */
int o_aak_phase(const double *pars, double T_fit, int length, double dt, int backint, double *out) {
  const double iota = pars[0], s = pars[1], p = pars[2], e = pars[3], M = pars[4], mu = pars[5];
  const double gamma = pars[6], psi = pars[7], alph = pars[8], qS = pars[9], phiS = pars[10];
  const double qK = pars[11], phiK = pars[12];
  double e_traj[100], v_map[100], M_map[100], s_map[100], dt_map[100], Phi0 = 0;
  const int steps = o_sparse_stage(T_fit, length, dt, iota, s, p, e, M, mu, psi,
                                   e_traj, v_map, M_map, s_map, dt_map, &Phi0, NULL);
  if (steps < 3) return steps < 0 ? steps : -2;
  double par[12] = {mu * O_MSUN_SEC, M_map[0] * O_MSUN_SEC, s_map[0], e_traj[0], iota, gamma, Phi0,
                    qS, phiS, qK, phiK, alph};
  const int j0 = o_pn_j0(dt, length, dt_map, steps);
  if (j0 < steps || j0 > 400000000) { o_set_error("knot count out of range"); return -3; }
  const int cap = 2 * j0 + 8;
  double *kn = malloc(sizeof(double) * 5 * (size_t)cap);
  int info[9];
  const int K = o_pn_knots(dt, length, par, e_traj, v_map, M, M_map, s, s_map, dt_map, steps, backint,
                           kn, kn + cap, kn + 2 * cap, kn + 3 * cap, kn + 4 * cap, cap, info);
  if (K < 0) { free(kn); return K; }
  const size_t L = (size_t)length + 1;
  double *dn = malloc(sizeof(double) * 12 * L);
  double *Pd = dn + 10 * L, *Ad = dn + 11 * L;
  const double t_fit_jmin = kn[info[2] - info[5]];
  int i_plunge = 0, i_buffer = 0;
  const int rc = o_pn_dense_rates(dt, length, par, backint, kn, kn + cap, kn + 2 * cap, kn + 3 * cap,
                                  kn + 4 * cap, K, info[0], info[7], t_fit_jmin, dn, &i_plunge, &i_buffer, Pd, Ad);
  if (rc == 0) {
    const double *t = dn, *ev = dn + L, *gim = dn + 5 * L, *Phi = dn + 6 * L, *alp = dn + 7 * L, *gd = dn + 9 * L;
    for (int i = 0; i < length; i++) {
      out[i] = t[i];
      out[length + i] = Phi[i];
      out[2 * (size_t)length + i] = Phi[i] + gim[i];
      out[3 * (size_t)length + i] = Phi[i] + gim[i] + alp[i];
      out[4 * (size_t)length + i] = Pd[i];
      out[5 * (size_t)length + i] = Pd[i] + gd[i];
      out[6 * (size_t)length + i] = Pd[i] + gd[i] + Ad[i];
      out[7 * (size_t)length + i] = ev[i];
    }
  }
  free(dn);
  free(kn);
  return rc;
}
