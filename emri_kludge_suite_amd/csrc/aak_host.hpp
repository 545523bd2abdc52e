// aak_host.hpp -- host half of the AAK waveform: sparse inspiral, frequency
// map, AK evolution/fit/plunge (knots) and the dense-stage parameters.
#pragma once

#include <complex>
#include <string>
#include <vector>

#include "aak_dense.hpp"
#include "aak_mlp.hpp"

namespace aak {

using HostCx = std::complex<double>;

struct WaveConfig {
  double T_fit = 1.0;
  int length = 0;
  double dt = 0.0;
  int LISA = 1;
  int backint = 0;
  int n_max = 0;  // > 0 overrides the reference harmonic rule
  int mixed = 0;  // 1: mixed-precision mode sum
  // mode-sum expansion blocks of 256 << exp_shift samples; -1: 2 in the
  // mixed form, 0 in FP64 (any value keeps the accuracy: the Taylor order
  // adapts, and blocks beyond order 8 take the per-sample path)
  int exp_shift = -1;
  // start point of the frequency map: 0 the reference's (1/sqrt(p), M, s),
  // 1 the learned model `mlp` plus a Newton polish (hybrid solver as fallback)
  int map_init = 0;
  aak_mlp_s *mlp = nullptr;
};

// 14 source parameters in the reference order (AAK_manager.hh:134-148).
struct SourceParams {
  double iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D;
};

struct SparseTrajectory {
  int steps = 0;
  double e_traj[100], v_map[100], M_map[100], s_map[100], dt_map[100];
  double Phi0 = 0.0;
  int map_iters[100];  // hybrid iterations, or Newton steps of the learned start (negative: fell back)
  int map_fallbacks = 0;
};

// Knots of the AK evolution.  The arrays span the whole index range j of the
// evolution (2*j0 entries, reused between waveforms); the K knots handed to
// the spline are entries [off, off+K), off = j_start.
struct KnotSet {
  std::vector<double> t, e, v, M, S;
  std::vector<double> eA, vA;  // raw AK evolution (before the fit corrections)
  int off = 0;
  int K = 0, i0 = 0, j0 = 0, j_min = 0, j_max = 0, j_max_temp = 0, j_start = 0, j_end = 0, plunged = 0, points = 0;
  int steps = 0;               // knots j0 .. j0+steps-1 take their times from dt_map
  double dt_large = 0.0;       // knot spacing outside the dt_map part
  double t_fit_jmin = 0.0;
  const double *kt() const { return t.data() + off; }
  const double *ky(int a) const {
    const std::vector<double> *arr[4] = {&e, &v, &M, &S};
    return arr[a]->data() + off;
  }
};

// Device knot table (kernels read it as uploaded):
//   [0, 4K)        knot values y of e, v, M, S
//   [4K, 8K)       natural-spline second-derivative coefficients c (y'' / 2)
//   [8K, 8K+H)     knot times of the irregular head (H = head knots), see
//                  KnotGrid for the regular part

struct WaveformPlan {
  SparseTrajectory traj;
  KnotSet knots;
  std::vector<double> table;  // device knot table when no external buffer is given
  double *tab = nullptr;      // the table actually filled (table.data() or external)
  size_t tab_n = 0;
  DenseParams P;
  double par[12];
};

// Supplies the storage of a device knot table (e.g. pinned host memory).
struct TableSink {
  virtual double *reserve(size_t n) = 0;
  virtual ~TableSink() {}
};

int sparse_stage(const WaveConfig &cfg, const SourceParams &sp, SparseTrajectory *out, std::string *err);
int build_knots(const WaveConfig &cfg, const double *par, double M_phys, double S_phys, const SparseTrajectory &tr,
                KnotSet *ks, std::string *err);
// Grow the evolution arrays of ks to L entries (no-op when they are larger).
void reserve_knots(KnotSet *ks, size_t L);
// Natural cubic splines of four arrays on one abscissa (one factorisation).
// y_out (optional) receives a copy of the four arrays.  With flat_from the
// constant tail is not written: *flat_from = first knot from which every
// interval is flat (equal values, zero coefficients), n - 1 if none.
int spline_natural4(const double *x, const double *const *ys, int n, double *const *cs,
                    double *const *y_out = nullptr, int *flat_from = nullptr);
int spline_natural(const double *x, const double *y, int n, double *c);
size_t table_size(const KnotSet &ks);
void fill_table(const KnotSet &ks, const SparseTrajectory &tr, double *tab, KnotGrid *grid);
double table_eval(const double *tab, const KnotGrid &g, int array, double t);
// Plunge sample ip and taper length ib (ten azimuthal periods at ip, + 1)
// from the knots and their device table (AAK.cc:240-242); *tp = ip's time.
int plunge_buffer(const WaveConfig &cfg, const KnotSet &ks, const double *tab, const KnotGrid &grid, double coslam,
                  int *ip, int *ib, double *tp, std::string *err);
int plan_waveform(const WaveConfig &cfg, const SourceParams &sp, WaveformPlan *plan, std::string *err,
                  TableSink *sink = nullptr);
// For a plan whose device stage skips the samples past the signal
// (DenseParams::full_dense = 0, as in likelihood batches): the knots past
// the interval of the last sample that stage evaluates (the rest of the
// last rates block of kRatesBlock samples after last_fold + 1) are dropped
// from the table, in place.  Every evaluated sample reads the same interval
// and values, so the results are unchanged.
void trim_table_to_signal(WaveformPlan *plan);
double omega_phi(double v, double e, double cosiota, double s, double M);

}  // namespace aak
