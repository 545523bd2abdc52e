// aak_host.hpp -- host half of the AAK waveform: sparse inspiral, frequency
// map, AK evolution/fit/plunge (knots) and the dense-stage parameters.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "aak_dense.hpp"

namespace aak {

using HostCx = std::complex<double>;

struct WaveConfig {
  double T_fit = 1.0;
  int length = 0;
  double dt = 0.0;
  int LISA = 1;
  int backint = 0;
  int n_max = 0;  // > 0 overrides the reference harmonic rule
};

// 14 source parameters in the reference order (AAK_manager.hh:134-148).
struct SourceParams {
  double iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D;
};

struct SparseTrajectory {
  int steps = 0;
  double e_traj[100], v_map[100], M_map[100], s_map[100], dt_map[100];
  double Phi0 = 0.0;
  int map_iters[100];
};

struct KnotSet {
  std::vector<double> t, e, v, M, S;
  int K = 0, i0 = 0, j0 = 0, j_min = 0, j_max = 0, j_max_temp = 0, j_start = 0, j_end = 0, plunged = 0, points = 0;
  double t_fit_jmin = 0.0;
};

struct WaveformPlan {
  SparseTrajectory traj;
  KnotSet knots;
  std::vector<double> packed;  // 17*K spline table for the device
  DenseParams P;
  double par[12];
};

int sparse_stage(const WaveConfig &cfg, const SourceParams &sp, SparseTrajectory *out, std::string *err);
int build_knots(const WaveConfig &cfg, const double *par, double M_phys, double S_phys, const SparseTrajectory &tr,
                KnotSet *ks, std::string *err);
int spline_natural(const double *x, const double *y, int n, double *c);
void pack_spline(const KnotSet &ks, std::vector<double> *packed);
double packed_eval(const std::vector<double> &packed, int K, int array, double t);
int plan_waveform(const WaveConfig &cfg, const SourceParams &sp, WaveformPlan *plan, std::string *err);
double omega_phi(double v, double e, double cosiota, double s, double M);

}  // namespace aak
