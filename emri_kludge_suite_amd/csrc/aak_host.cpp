// aak_host.cpp -- host half of the AAK waveform.
//
//   sparse_stage   AAKpy.cc:20-44 / AAK_manager.cu:296-320: inspiral, Kerr
//                  frequencies, mean anomaly, frequency map at <= 100 points
//   build_knots    AAK.cc:27-226: AK evolution, quadratic fit corrections,
//                  plunge search and the spline knots
//   plan_waveform  knots -> natural-spline table, plunge index/buffer and the
//                  per-waveform constants of the device mode sum
//
// Compiled without floating-point contraction: the host stage reproduces the
// reference CPU AAK's rounding, so the device fold can be exact.
#include "aak_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>

#include "aak_bessel.hpp"
#include "aak_freqmap.hpp"
#include "aak_inspiral.hpp"

namespace aak {

double omega_phi(double v, double e, double cosiota, double s, double M) { return PN::omega_phi(v, e, cosiota, s, M); }

int sparse_stage(const WaveConfig &cfg, const SourceParams &sp, SparseTrajectory *out, std::string *err) {
  const int maxsteps = 100;
  Inspiral<HostCx> insp;
  if (insp.init(sp.s)) {
    *err = "LSO table failed";
    return -1;
  }
  const double tr = cfg.length * cfg.dt / kMsunSec / sp.M / sp.M * sp.mu;
  const double dt_fit = ((tr < cfg.T_fit) ? tr : cfg.T_fit) / (maxsteps - 1);
  InspiralPoint pts[100];
  const int steps = insp.eccentric(sp.p, sp.e, cos(sp.iota), dt_fit, maxsteps, pts);
  if (steps < 0) {
    *err = "inspiral integration failed";
    return -1;
  }
  if (steps < 3) {
    *err = "trajectory too short (p0 at or inside the LSO?)";
    return -2;
  }
  // learned start points of all maps at once (one batched inference)
  const bool learned = cfg.map_init == 1 && cfg.mlp;
  double starts[3 * 100];
  if (learned) {
    double q[4 * 100];
    for (int i = 0; i < steps; i++) {
      q[4 * i] = sp.s;
      q[4 * i + 1] = pts[i].p;
      q[4 * i + 2] = pts[i].e;
      q[4 * i + 3] = sp.iota;
    }
    if (mlp_infer(cfg.mlp, q, steps, starts) != 0) {
      *err = "frequency-map model inference failed";
      return -1;
    }
  }
  out->map_fallbacks = 0;
  for (int i = 0; i < steps; i++) {
    Geodesic<HostCx> g;
    if (g.init(pts[i].p, pts[i].e, pts[i].cosi, sp.s, 1)) {
      *err = "geodesic constants failed";
      return -1;
    }
    double Om[3];
    g.frequencies(Om);
    if (i == 0) {
      const double e = sp.e, psi = sp.psi;
      out->Phi0 = atan2(sqrt(1. - e * e) * sin(psi), e + cos(psi)) - e * sqrt(1. - e * e) * sin(psi) / (1. + e * cos(psi));
    }
    FreqMap fm{Om, sp.M, pts[i].e, sp.iota};
    double x[3] = {1. / sqrt(pts[i].p), sp.M, sp.s};
    if (learned) {
      // Newton polish of the learned start to the reference's tolerance (or
      // 1e-12 M, whichever is smaller); the reference's solve otherwise
      const double y0[3] = {starts[3 * i], sp.M * starts[3 * i + 1], starts[3 * i + 2]};
      double y[3] = {y0[0], y0[1], y0[2]};
      const int k = fm.polish(y, fmin(1.e-6, 1.e-12 * sp.M), 8);
      if (k >= 0 && FreqMap::near_start(y, y0, 0.05)) {
        x[0] = y[0];
        x[1] = y[1];
        x[2] = y[2];
        out->map_iters[i] = k;
      } else {
        out->map_iters[i] = -fm.solve(x);
        out->map_fallbacks++;
      }
    } else {
      out->map_iters[i] = fm.solve(x);
    }
    out->e_traj[i] = pts[i].e;
    out->v_map[i] = x[0];
    out->M_map[i] = x[1];
    out->s_map[i] = x[2];
    out->dt_map[i] = pts[i].t * kMsunSec * sp.M * sp.M / sp.mu;
  }
  out->steps = steps;
  return steps;
}

// y - y0 = c0 (x - x0) + c1 (x - x0)^2 by least squares: column-scaled
// Householder QR of the (n-1) x 2 design matrix.
static int polyfit2(double *coeff, const double *x, const double *y, int n) {
  const int m = n - 1;
  if (m < 2) return -1;
  std::vector<double> A0(m), A1(m), b(m);
  for (int i = 0; i < m; i++) {
    A0[i] = x[i + 1] - x[0];
    A1[i] = pow(x[i + 1] - x[0], 2);
    b[i] = y[i + 1] - y[0];
  }
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < m; i++) {
    s0 += A0[i] * A0[i];
    s1 += A1[i] * A1[i];
  }
  s0 = sqrt(s0);
  s1 = sqrt(s1);
  if (s0 == 0.0) s0 = 1.0;
  if (s1 == 0.0) s1 = 1.0;
  for (int i = 0; i < m; i++) {
    A0[i] /= s0;
    A1[i] /= s1;
  }
  double *col[2] = {A0.data(), A1.data()};
  double diag[2];
  for (int k = 0; k < 2; k++) {
    double nrm = 0.0;
    for (int i = k; i < m; i++) nrm += col[k][i] * col[k][i];
    nrm = sqrt(nrm);
    const double alpha = (col[k][k] > 0) ? -nrm : nrm;
    const double vk = col[k][k] - alpha;
    double vtv = vk * vk;
    for (int i = k + 1; i < m; i++) vtv += col[k][i] * col[k][i];
    if (vtv != 0.0) {
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
    diag[k] = alpha;
  }
  const double z1 = b[1] / diag[1];
  const double z0 = (b[0] - col[1][0] * z1) / diag[0];
  coeff[0] = z0 / s0;
  coeff[1] = z1 / s1;
  return 0;
}

static double pn_p(double v, double e, double coslam, double S, double M, double Mphys) {
  return (1. - e * e) * pow(PN::omega_phi(v, e, coslam, S, M) * Mphys * kMsunSec, -2. / 3.);
}

static bool same_bits(double a, double b) {
  uint64_t x, y;
  std::memcpy(&x, &a, sizeof x);
  std::memcpy(&y, &b, sizeof y);
  return x == y;
}

void reserve_knots(KnotSet *ks, size_t L) {
  for (std::vector<double> *a : {&ks->eA, &ks->vA, &ks->t, &ks->e, &ks->v, &ks->M, &ks->S})
    if (a->size() < L) a->resize(L);
}

int build_knots(const WaveConfig &cfg, const double *par, double M_phys, double S_phys, const SparseTrajectory &tr,
                KnotSet *ks, std::string *err) {
  const double timestep = cfg.dt;
  const int vlength = cfg.length, steps = tr.steps;
  const double mu = par[0], M0 = par[1], S0 = par[2], e0 = par[3];
  const double coslam = cos(par[4]);
  const double *dt_map = tr.dt_map;
  const double dt_large = dt_map[1];
  const int j0 = (int)((vlength * timestep - dt_map[steps - 1]) / dt_large + steps + 1);
  if (j0 < steps || j0 > 400000000) {
    *err = "knot count out of range";
    return -3;
  }
  int j_start = j0, j_min = j0, j_end = 2 * j0 - 1, j_max = j_end, j_max_temp = j_max;
  double t_end = vlength * timestep;
  // Arrays over the whole index range j in [0, 2 j0); every entry is written
  // before it is read, and storage is reused across waveforms.
  const size_t L = (size_t)2 * j0;
  reserve_knots(ks, L);
  double *eA = ks->eA.data(), *vA = ks->vA.data(), *tf = ks->t.data(), *ef = ks->e.data(), *vf = ks->v.data(),
         *Mf = ks->M.data(), *Sf = ks->S.data();
  double edot = 0, vdot = 0, edotm, vdotm;
  eA[j0] = e0;
  vA[j0] = tr.v_map[0];
  auto clamp_step = [&](int j) {
    if (eA[j + 1] < 0. || vA[j + 1] < vA[j] || std::isnan(eA[j + 1]) || std::isnan(vA[j + 1])) {
      if (j_max == j_max_temp) j_max = j;
      eA[j + 1] = eA[j_max];
      vA[j + 1] = vA[j_max];
      return true;
    }
    return false;
  };
  for (int j = j0; j < j0 + steps - 1; j++) {
    edotm = edot;
    vdotm = vdot;
    PN::ak_rates(vA[j], eA[j], coslam, mu, M0, S0, &edot, &vdot);
    if (j == j0) {
      edotm = edot;
      vdotm = vdot;
    }
    eA[j + 1] = eA[j] + (1.5 * edot - .5 * edotm) * (dt_map[j - j0 + 1] - dt_map[j - j0]);
    vA[j + 1] = vA[j] + (1.5 * vdot - .5 * vdotm) * (dt_map[j - j0 + 1] - dt_map[j - j0]);
    clamp_step(j);
  }
  j_max_temp = j_max;
  int points = std::min(steps, j_max - j0 + 1);
  while (points >= 2 && tr.e_traj[points - 1] > tr.e_traj[points - 2]) points--;
  if (points < 3) {
    *err = "Fitting error: t_0 too close to plunge";
    return -4;
  }
  double ec[2], vc[2], Mc[2], Sc[2];
  {
    std::vector<double> ediff(points), vdiff(points);
    for (int i = 0; i < points; i++) {
      ediff[i] = tr.e_traj[i] - eA[i + j0];
      vdiff[i] = tr.v_map[i] - vA[i + j0];
    }
    polyfit2(ec, dt_map, ediff.data(), points);
    polyfit2(vc, dt_map, vdiff.data(), points);
    polyfit2(Mc, dt_map, tr.M_map, points);
    polyfit2(Sc, dt_map, tr.s_map, points);
  }
  // Clamp cycles.  After this loop's first clamp, j_max is fixed and every
  // clamp at a step j resets (e, v) to the same state (eA[j_max], vA[j_max]);
  // step j + 1 then advances from that state with the rates of step j as its
  // previous rates.  So the steps after a clamp depend only on the bits of
  // (edot, vdot) at the clamped step: when a clamp's rates equal those of an
  // earlier clamp at step j', every later entry repeats the one j - j' before
  // it (fixed step, pure rates, the same clamp decisions).  Sources that plunge
  // long before the end of the evolution settle into such a cycle within a
  // few clamps; the rates of the clamps seen so far are kept in a small
  // open-addressed table.
  constexpr int kCycleSlots = 1 << 12, kCycleMax = kCycleSlots / 2;
  thread_local std::vector<std::pair<uint64_t, uint64_t>> cyc_key;
  thread_local std::vector<int> cyc_j;
  int cyc_n = 0;
  bool cyc_on = false;
  for (int j = j0 + steps - 1; j < j_end; j++) {
    edotm = edot;
    vdotm = vdot;
    PN::ak_rates(vA[j], eA[j], coslam, mu, M0, S0, &edot, &vdot);
    eA[j + 1] = eA[j] + (1.5 * edot - .5 * edotm) * dt_large;
    vA[j + 1] = vA[j] + (1.5 * vdot - .5 * vdotm) * dt_large;
    const bool clamped = clamp_step(j);
    // Past a plunge the clamp resets (e, v) to the same point at every step:
    // once a step starts from the state it produces, with the rates it had
    // before, every later step repeats it exactly (fixed step, pure rates), so
    // the remaining entries are that state.
    if (same_bits(eA[j + 1], eA[j]) && same_bits(vA[j + 1], vA[j]) && same_bits(edot, edotm) && same_bits(vdot, vdotm)) {
      for (int k = j + 2; k <= j_end; k++) {
        eA[k] = eA[j + 1];
        vA[k] = vA[j + 1];
      }
      break;
    }
    if (clamped && j_max != j_max_temp && cyc_n < kCycleMax) {
      if (!cyc_on) {
        cyc_key.assign(kCycleSlots, {0, 0});
        cyc_j.assign(kCycleSlots, -1);
        cyc_on = true;
      }
      uint64_t ke, kv;
      std::memcpy(&ke, &edot, sizeof ke);
      std::memcpy(&kv, &vdot, sizeof kv);
      unsigned h = (unsigned)((ke * 0x9E3779B97F4A7C15ULL ^ kv) >> 40) & (kCycleSlots - 1);
      while (cyc_j[h] >= 0 && !(cyc_key[h].first == ke && cyc_key[h].second == kv)) h = (h + 1) & (kCycleSlots - 1);
      if (cyc_j[h] >= 0) {
        const int period = j - cyc_j[h];
        for (int k = j + 2; k <= j_end; k++) {
          eA[k] = eA[k - period];
          vA[k] = vA[k - period];
        }
        break;
      }
      cyc_key[h] = {ke, kv};
      cyc_j[h] = j;
      cyc_n++;
    }
  }
  j_max_temp = j_max;
  {
    // Knot values.  Each knot's fitted values depend on that knot alone, so
    // they are formed for every knot first (loops without branches, which
    // vectorise); the clamp then runs in knot order from the first knot it
    // touches, on the same values: before that knot it changes nothing.
    double *__restrict tq = tf, *__restrict eq = ef, *__restrict vq = vf, *__restrict Mq = Mf, *__restrict Sq = Sf;
    const double *__restrict eAq = eA, *__restrict vAq = vA;
    const double ec0 = ec[0], ec1 = ec[1], vc0 = vc[0], vc1 = vc[1], Mc0 = Mc[0], Mc1 = Mc[1], Sc0 = Sc[0], Sc1 = Sc[1];
    const int jh = std::min(j0 + steps - 1, j_end);  // last knot of the sparse head
    for (int j = j0; j <= jh; j++) {
      const double dt = dt_map[j - j0], dt2 = dt * dt;
      tq[j] = dt;
      eq[j] = eAq[j] + ec0 * dt + ec1 * dt2;
      vq[j] = vAq[j] + vc0 * dt + vc1 * dt2;
      Mq[j] = M0 + (Mc0 * dt + Mc1 * dt2) * kMsunSec;
      Sq[j] = S0 + Sc0 * dt + Sc1 * dt2;
    }
    const double th = dt_map[steps - 1];
    const int jb = j0 + steps - 1;
    for (int j = jh + 1; j <= j_end; j++) {
      const double dt = th + dt_large * (double)(j - jb), dt2 = dt * dt;
      tq[j] = dt;
      eq[j] = eAq[j] + ec0 * dt + ec1 * dt2;
      vq[j] = vAq[j] + vc0 * dt + vc1 * dt2;
      Mq[j] = M0 + (Mc0 * dt + Mc1 * dt2) * kMsunSec;
      Sq[j] = S0 + Sc0 * dt + Sc1 * dt2;
    }
    int j = j0;
    while (j <= j_end && !(ef[j] < 0. || vf[j] < vf[std::max(j0, j - 1)])) j++;
    for (; j <= j_end && j_max == j_max_temp; j++) {
      if (ef[j] < 0. || vf[j] < vf[std::max(j0, j - 1)]) {
        j_max = j - 1;
        ef[j] = ef[j_max];
        vf[j] = vf[j_max];
        Mf[j] = Mf[j_max];
        Sf[j] = Sf[j_max];
      }
    }
    // j_max is fixed from here on, so is the clamp state: selects instead of
    // branches (the clamps past a plunge follow no pattern a predictor learns)
    if (j <= j_end) {
      const double Ec = ef[j_max], Vc = vf[j_max], Mc_ = Mf[j_max], Sc_ = Sf[j_max];
      double vprev = vf[std::max(j0, j - 1)];
      for (; j <= j_end; j++) {
        const bool c = ef[j] < 0. || vf[j] < vprev;
        ef[j] = c ? Ec : ef[j];
        vf[j] = c ? Vc : vf[j];
        Mf[j] = c ? Mc_ : Mf[j];
        Sf[j] = c ? Sc_ : Sf[j];
        vprev = vf[j];
      }
    }
  }
  j_max_temp = j_max;
  const int j_RR = (int)((kMsunSec * M_phys * kMsunSec * M_phys / mu) / dt_large) + 1;
  const double z1 = 1. + pow(1. - S_phys * S_phys, 1. / 3.) * (pow(1. + S_phys, 1. / 3.) + pow(1. - S_phys, 1. / 3.));
  const double z2 = sqrt(3. * S_phys * S_phys + z1 * z1);
  const double LSO_min = 3. + z2 - sqrt((3. - z1) * (3. + z1 + 2. * z2));
  const double LSO_max = 3. + z2 + sqrt((3. - z1) * (3. + z1 + 2. * z2));
  auto unstable = [&](int j, int *bad) {
    Geodesic<HostCx> g;
    if (g.init(pn_p(vf[j], ef[j], coslam, Sf[j], Mf[j], M_phys), ef[j], coslam, S_phys, 1)) {
      *bad = 1;
      return false;
    }
    return g.stable == -1 || g.E > 1;
  };
  int bad = 0;
  // p_j = (1 - e^2) x^(-2/3) with x = Omega_phi M_phys M_sun (pn_p).  Where
  // (1 - e^2)^3 >= LSO_max^3 (1 + 1e-9) x^2, p_j >= LSO_max for certain (the
  // margin dwarfs every rounding), so neither test below can pass and the pow
  // is skipped; elsewhere p_j is formed exactly as pn_p forms it.
  const double L3 = LSO_max * LSO_max * LSO_max * (1.0 + 1e-9);
  // x and the skip test of a chunk of knots first (independent per knot, so
  // the loop vectorises), then the scan itself in knot order
  constexpr int kScanChunk = 64;
  double xs[kScanChunk];
  bool skip[kScanChunk];
  bool scan_done = false;
  for (int jc = j0; jc <= j_max_temp && !scan_done; jc += kScanChunk) {
    const int m = std::min(kScanChunk, j_max_temp - jc + 1);
    // PN::omega_phi with its orientation chosen once: (cos i, s) when
    // cos i > 0, else (-cos i, -s)
    const bool pro = coslam > 0;
    const double Y = pro ? coslam : -coslam;
    for (int u = 0; u < m; u++) {
      const int j = jc + u;
      const double q = pro ? Sf[j] : -Sf[j];
      const double x = PN::dphidm(vf[j], ef[j], Y, q) / PN::dtdm(vf[j], ef[j], Y, q) / Mf[j] * M_phys * kMsunSec;
      const double ome = 1. - ef[j] * ef[j];
      xs[u] = x;
      skip[u] = (x > 0.) & (ome > 0.) & (ome * ome * ome >= L3 * (x * x));
    }
    for (int u = 0; u < m; u++) {
      const int j = jc + u;
      if (j > j_max) {
        scan_done = true;
        break;
      }
      if (skip[u]) continue;
      const double ome = 1. - ef[j] * ef[j];
      const double pj = ome * pow(xs[u], -2. / 3.);
      if (pj < LSO_min && j_max == j_max_temp) {
        j_min = std::max(j - j_RR, j0);
        j_max = std::min(j + 1, j_max_temp);
      }
      if (j > j0 && (j - j0) % j_RR == 0 && pj < LSO_max && j_max == j_max_temp) {
        j_min = std::max(j - j_RR, j0);
        if (unstable(j, &bad)) j_max = std::min(j + 1, j_max_temp);
        if (bad) {
          scan_done = true;
          break;
        }
      }
    }
  }
  while (!bad && j_max - j_min > 1) {
    const int jm = (j_max + j_min) / 2;
    if (unstable(jm, &bad)) j_max = jm;
    else j_min = jm;
  }
  if (bad) {
    *err = "geodesic constants failed in the plunge search";
    return -5;
  }
  if (cfg.backint) {
    double edotp, vdotp;
    j_end = j_max;
    t_end = (tf[j_end] < vlength * timestep) ? tf[j_end] : vlength * timestep;
    j_start = j0 + floor((t_end - vlength * timestep) / dt_large);
    for (int j = j0; j > j_start; j--) {
      edotp = edot;
      vdotp = vdot;
      PN::ak_rates(vA[j], eA[j], coslam, mu, M0, S0, &edot, &vdot);
      if (j == j0) {
        edotp = edot;
        vdotp = vdot;
      }
      eA[j - 1] = eA[j] - (1.5 * edot - .5 * edotp) * dt_large;
      vA[j - 1] = vA[j] - (1.5 * vdot - .5 * vdotp) * dt_large;
    }
    for (int j = j_start; j < j0; j++) {
      const double dtu = dt_large * (j - j0);
      double dt;
      if (j > j0 - points) dt = (dtu + dtu * (j - j0 + points) / points) / 2.;
      else dt = -dt_large * points / 2.;
      const double dt2 = dt * dt;
      tf[j] = dtu;
      ef[j] = eA[j] + ec[0] * dt + ec[1] * dt2;
      vf[j] = vA[j] + vc[0] * dt + vc[1] * dt2;
      Mf[j] = M0 + (Mc[0] * dt + Mc[1] * dt2) * kMsunSec;
      Sf[j] = S0 + Sc[0] * dt + Sc[1] * dt2;
    }
  }
  const int K = j_end - j_start + 1;
  ks->K = K;
  ks->off = j_start;
  ks->steps = steps;
  ks->dt_large = dt_large;
  ks->i0 = vlength - (int)(t_end / timestep);
  ks->j0 = j0;
  ks->j_min = j_min;
  ks->j_max = j_max;
  ks->j_max_temp = j_max_temp;
  ks->j_start = j_start;
  ks->j_end = j_end;
  ks->plunged = (j_max != j_max_temp);
  ks->points = points;
  ks->t_fit_jmin = tf[j_min];
  return K;
}

// Natural cubic spline: symmetric tridiagonal system for c = y''/2 solved by
// LDL^T (the scheme of GSL's cspline, which the reference links).  The
// factorisation depends on the abscissae only, so the four arrays of one
// waveform share it; each array's arithmetic is unchanged.  One forward pass
// runs the factorisation, the four forward substitutions and the diagonal
// scaling side by side (independent dependency chains), reusing each slope
// (y[i+1] - y[i]) / h_i for the two right-hand sides that contain it; the
// backward pass follows.  y_out, if given, receives a copy of the four arrays.
//
// A tail where all four arrays are constant (a source that plunged: the knots
// past the plunge hold the clamped state) has right-hand sides exactly +0, so
// the forward values there decay geometrically and become exactly +0 within a
// few hundred rows; from that row on every coefficient is +0 and the backward
// pass leaves the earlier ones unchanged (c - gam * (+0) = c), so the solve
// stops there and fills the rest.
int spline_natural4(const double *x, const double *const *ys, int n, double *const *cs, double *const *y_out,
                    int *flat_from) {
  if (n < 3) return -1;
  const int sys = n - 2;
  // first index of the tail on which every array equals its last value, bit for bit
  int t_const = n - 1;
  {
    uint64_t last[4];
    for (int a = 0; a < 4; a++) std::memcpy(&last[a], &ys[a][n - 1], sizeof(uint64_t));
    while (t_const > 0) {
      bool same = true;
      for (int a = 0; a < 4 && same; a++) {
        uint64_t u;
        std::memcpy(&u, &ys[a][t_const - 1], sizeof(uint64_t));
        same = (u == last[a]);
      }
      if (!same) break;
      t_const--;
    }
  }
  thread_local std::vector<double> gam_v;
  if ((int)gam_v.size() < sys) gam_v.resize(sys);
  double *gam = gam_v.data();
  const double *y0 = ys[0], *y1 = ys[1], *y2 = ys[2], *y3 = ys[3];
  double *c0 = cs[0], *c1 = cs[1], *c2 = cs[2], *c3 = cs[3];
  for (int a = 0; a < 4; a++) {
    cs[a][0] = 0.0;
    cs[a][n - 1] = 0.0;
  }
  if (y_out)
    for (int a = 0; a < 4; a++) {
      y_out[a][0] = ys[a][0];
      y_out[a][1] = ys[a][1];
    }
  // h_i = x[i+1] - x[i], gi_i = 1/h_i (0 for h_i = 0); the right-hand side of
  // row i is g_i = 3 ((y[i+2] - y[i+1]) gi_{i+1} - (y[i+1] - y[i]) gi_i)
  double h0 = x[1] - x[0];
  double gi0 = (h0 != 0.0) ? 1.0 / h0 : 0.0;
  double s0 = (y0[1] - y0[0]) * gi0, s1 = (y1[1] - y1[0]) * gi0;
  double s2 = (y2[1] - y2[0]) * gi0, s3 = (y3[1] - y3[0]) * gi0;
  double z0 = 0.0, z1 = 0.0, z2 = 0.0, z3 = 0.0, od_prev = 0.0, gam_prev = 0.0;
  int stop = sys;  // rows [stop, sys) are +0
  for (int i = 0; i < sys; i++) {
    const double h1 = x[i + 2] - x[i + 1];
    const double gi1 = (h1 != 0.0) ? 1.0 / h1 : 0.0;
    const double od = h1, dg = 2.0 * (h1 + h0);
    const double al = (i == 0) ? dg : dg - od_prev * gam_prev;
    const double t0 = (y0[i + 2] - y0[i + 1]) * gi1, t1 = (y1[i + 2] - y1[i + 1]) * gi1;
    const double t2 = (y2[i + 2] - y2[i + 1]) * gi1, t3 = (y3[i + 2] - y3[i + 1]) * gi1;
    const double g0 = 3.0 * (t0 - s0), g1 = 3.0 * (t1 - s1), g2 = 3.0 * (t2 - s2), g3 = 3.0 * (t3 - s3);
    if (i == 0) {
      z0 = g0;
      z1 = g1;
      z2 = g2;
      z3 = g3;
    } else {
      z0 = g0 - gam_prev * z0;
      z1 = g1 - gam_prev * z1;
      z2 = g2 - gam_prev * z2;
      z3 = g3 - gam_prev * z3;
    }
    c0[1 + i] = z0 / al;
    c1[1 + i] = z1 / al;
    c2[1 + i] = z2 / al;
    c3[1 + i] = z3 / al;
    if (y_out) {
      y_out[0][i + 2] = y0[i + 2];
      y_out[1][i + 2] = y1[i + 2];
      y_out[2][i + 2] = y2[i + 2];
      y_out[3][i + 2] = y3[i + 2];
    }
    if (i < sys - 1) {
      gam_prev = od / al;
      gam[i] = gam_prev;
    }
    od_prev = od;
    h0 = h1;
    s0 = t0;
    s1 = t1;
    s2 = t2;
    s3 = t3;
    if (i >= t_const && z0 == 0.0 && z1 == 0.0 && z2 == 0.0 && z3 == 0.0) {
      stop = i + 1;
      break;
    }
  }
  // knots from J on bound only flat intervals (equal values, zero coefficients)
  const int J = (stop < sys) ? std::max(stop + 1, t_const) : n - 1;
  if (flat_from) *flat_from = J;
  const int fill_end = flat_from ? J + 1 : n;  // the caller keeps knots [0, J] only
  if (stop < sys) {
    for (int a = 0; a < 4; a++) {
      for (int m = stop + 1; m < std::min(fill_end, n - 1); m++) cs[a][m] = 0.0;
      if (y_out)
        for (int k = stop + 2; k < fill_end; k++) y_out[a][k] = ys[a][k];
    }
  }
  // c[1+i] = z_i / al_i - gam_i c[i+2]
  double n0 = c0[stop], n1 = c1[stop], n2 = c2[stop], n3 = c3[stop];
  for (int i = stop - 2; i >= 0; i--) {
    const double gm = gam[i];
    n0 = c0[1 + i] - gm * n0;
    n1 = c1[1 + i] - gm * n1;
    n2 = c2[1 + i] - gm * n2;
    n3 = c3[1 + i] - gm * n3;
    c0[1 + i] = n0;
    c1[1 + i] = n1;
    c2[1 + i] = n2;
    c3[1 + i] = n3;
  }
  return 0;
}

int spline_natural(const double *x, const double *y, int n, double *c) {
  thread_local std::vector<double> dummy;
  if ((int)dummy.size() < 3 * n) dummy.resize(3 * n);
  const double *ys[4] = {y, y, y, y};
  double *cs[4] = {c, dummy.data(), dummy.data() + n, dummy.data() + 2 * n};
  return spline_natural4(x, ys, n, cs);
}

size_t table_size(const KnotSet &ks) { return (size_t)kTableArrays * ks.K + ks.steps; }

void fill_table(const KnotSet &ks, const SparseTrajectory &tr, double *tab, KnotGrid *grid) {
  const int K = ks.K;
  const double *ys[4] = {ks.ky(0), ks.ky(1), ks.ky(2), ks.ky(3)};
  double *cs[4], *yo[4];
  for (int a = 0; a < 4; a++) {
    yo[a] = tab + (size_t)a * K;
    cs[a] = tab + (size_t)(4 + a) * K;
  }
  int J = K - 1;
  if (spline_natural4(ks.kt(), ys, K, cs, yo, &J) != 0) {
    for (int a = 0; a < 4; a++) std::memcpy(yo[a], ys[a], sizeof(double) * K);
    J = K - 1;
  }
  // A flat tail (from knot J on) is not stored: the device returns y[J] + 0 past
  // knot J, which is what the cubic of every flat interval evaluates to.
  const int Kt = std::max(J + 1, 2);
  if (Kt < K)
    for (int a = 1; a < kTableArrays; a++)
      std::memmove(tab + (size_t)a * Kt, tab + (size_t)a * K, sizeof(double) * Kt);
  std::memcpy(tab + (size_t)kTableArrays * Kt, tr.dt_map, sizeof(double) * ks.steps);
  grid->K = Kt;
  grid->head_off = ks.j0 - ks.j_start;
  grid->steps = ks.steps;
  grid->flat = (Kt < K) ? 1 : 0;
  grid->dt_large = ks.dt_large;
  grid->t_head_last = tr.dt_map[ks.steps - 1];
}

void trim_table_to_signal(WaveformPlan *plan) {
  DenseParams &P = plan->P;
  const long long last = ((long long)(P.last_fold + 1) / kRatesBlock) * kRatesBlock + kRatesBlock - 1;
  if (last >= P.N) return;  // sample N (time t_last, the last knot) is evaluated
  const int i_max = (int)std::max(last, 1LL);
  const KnotGrid g = P.grid;
  double *tab = plan->tab;
  const double *head = tab + (size_t)kTableArrays * g.K;
  const int Kn = knot_index(g, head, (double)(i_max - P.i0) * P.dt) + 2;
  if (Kn >= g.K) return;
  for (int a = 1; a < kTableArrays; a++) std::memmove(tab + (size_t)a * Kn, tab + (size_t)a * g.K, sizeof(double) * Kn);
  std::memmove(tab + (size_t)kTableArrays * Kn, head, sizeof(double) * g.steps);
  P.grid.K = Kn;
  P.grid.flat = 0;
  plan->tab_n = (size_t)kTableArrays * Kn + g.steps;
}

double table_eval(const double *tab, const KnotGrid &g, int array, double t) {
  const int K = g.K;
  const double *head = tab + (size_t)kTableArrays * K;
  if (g.flat && t >= knot_time(g, head, K - 1)) return tab[(size_t)array * K + K - 1] + 0.0;
  const int k = knot_index(g, head, t);
  const double x0 = knot_time(g, head, k), x1 = knot_time(g, head, k + 1);
  const double *y = tab + (size_t)array * K, *c = tab + (size_t)(4 + array) * K;
  double b, d;
  spline_interval(x0, x1, y[k], y[k + 1], c[k], c[k + 1], &b, &d);
  const double dl = t - x0;
  return y[k] + dl * (b + dl * (c[k] + dl * d));
}

int plunge_buffer(const WaveConfig &cfg, const KnotSet &ks, const double *tab, const KnotGrid &grid, double coslam,
                  int *ip_out, int *ib_out, double *tp_out, std::string *err) {
  const int N = cfg.length;
  const int ip = ks.plunged ? ks.i0 + (int)(ks.t_fit_jmin / cfg.dt) : N - 1;
  if (ip < 0 || ip > N) {
    *err = "plunge index outside the waveform";
    return -6;
  }
  // the dense grid: sample 0 at the first knot, sample N at the last, i - i0
  // steps of dt in between (AAK.cc:228-231)
  const double tp = ip == 0 ? ks.kt()[0] : (ip == N ? ks.kt()[ks.K - 1] : (double)(ip - ks.i0) * cfg.dt);
  const double ve = table_eval(tab, grid, 1, tp), ee = table_eval(tab, grid, 0, tp);
  const double Me = table_eval(tab, grid, 2, tp), Se = table_eval(tab, grid, 3, tp);
  *ib_out = (int)(10. / (PN::omega_phi(ve, ee, coslam, Se, Me) / 2. / kPi) / cfg.dt) + 1;
  *ip_out = ip;
  *tp_out = tp;
  return 0;
}

int plan_waveform(const WaveConfig &cfg, const SourceParams &sp, WaveformPlan *plan, std::string *err,
                  TableSink *sink) {
  int rc = sparse_stage(cfg, sp, &plan->traj, err);
  if (rc < 3) return rc < 0 ? rc : -2;
  const SparseTrajectory &tr = plan->traj;
  double *par = plan->par;
  par[0] = sp.mu * kMsunSec;
  par[1] = tr.M_map[0] * kMsunSec;
  par[2] = tr.s_map[0];
  par[3] = tr.e_traj[0];
  par[4] = sp.iota;
  par[5] = sp.gamma;
  par[6] = tr.Phi0;
  par[7] = sp.theta_S;
  par[8] = sp.phi_S;
  par[9] = sp.theta_K;
  par[10] = sp.phi_K;
  par[11] = sp.alph;
  int nmodes = (int)(30 * par[3]);
  if (par[3] < 0.135) nmodes = 4;
  if (cfg.n_max > 0) nmodes = cfg.n_max;
  rc = build_knots(cfg, par, sp.M, sp.s, tr, &plan->knots, err);
  if (rc < 0) return rc;
  const KnotSet &ks = plan->knots;
  DenseParams &P = plan->P;
  std::memset(&P, 0, sizeof(P));
  plan->tab_n = table_size(ks);
  if (sink) {
    plan->tab = sink->reserve(plan->tab_n);
    if (!plan->tab) {
      *err = "knot table storage unavailable";
      return -8;
    }
  } else {
    if (plan->table.size() < plan->tab_n) plan->table.resize(plan->tab_n);
    plan->tab = plan->table.data();
  }
  fill_table(ks, plan->traj, plan->tab, &P.grid);
  plan->tab_n = (size_t)kTableArrays * P.grid.K + ks.steps;
  const int N = cfg.length;
  P.N = N;
  P.i0 = ks.i0;
  P.K = ks.K;
  P.dt = cfg.dt;
  P.mich = cfg.LISA ? 1 : 0;
  P.nmodes = nmodes;
  P.t_first = ks.kt()[0];
  P.t_last = ks.kt()[ks.K - 1];
  const double coslam = cos(par[4]);
  P.coslam = coslam;
  P.x0[0] = par[5];
  P.x0[1] = par[6];
  P.x0[2] = par[11];
  // plunge index and taper length (AAK.cc:240-242)
  int ip = 0, ib = 0;
  double tp = 0.0;
  rc = plunge_buffer(cfg, ks, plan->tab, P.grid, coslam, &ip, &ib, &tp, err);
  if (rc < 0) return rc;
  P.ip = ip;
  P.ib = ib;
  P.last_fold = std::min(N - 1, ip + ib);
  if (P.last_fold < P.i0) P.last_fold = P.i0 - 1;
  P.guard_last = P.last_fold;
  P.t_plunge = tp;
  P.t_zero = tp + cfg.dt * ib;
  // mode-sum constants (AAK.cc:323-338, 434-479; KSParMap.cc:541-549)
  const double lam = par[4], qS = par[7], phiS = par[8], qK = par[9], phiK = par[10];
  P.lam = lam;
  P.sinlam = sin(lam);
  P.qS = qS;
  P.cosqS = cos(qS);
  P.sinqS = sin(qS);
  P.qK = qK;
  P.cosqK = cos(qK);
  P.sinqK = sin(qK);
  P.phiS = phiS;
  P.phiK = phiK;
  P.cosphiK = cos(phiK);
  P.sinphiK = sin(phiK);
  P.halfsqrt3 = sqrt(3.) / 2.;
  P.Sdotn = P.cosqK * P.cosqS + P.sinqK * P.sinqS * cos(phiK - phiS);
  P.sin_dphiKS = sin(phiK - phiS);
  P.cos_dphiKS = cos(phiK - phiS);
  P.psi_c0 = (P.cosqK * P.cosqS + P.sinqK * P.sinqS * cos(phiK - phiS));
  P.psi_c1 = .5 * P.sinqK * P.sinqS * sin(phiK - phiS);
  P.psi_c2 = (P.cosqK * P.sinqS * sin(phiS) - P.cosqS * P.sinqK * sin(phiK));
  P.psi_c3 = (P.cosqS * P.sinqK * cos(phiK) - P.cosqK * P.sinqS * cos(phiS));
  {
    const double up = (P.cosqS * P.sinqK * cos(phiS - phiK) - P.cosqK * P.sinqS);
    const double dw = (P.sinqK * sin(phiS - phiK));
    const double psi_ldc = (dw != 0.0) ? -atan2(up, dw) : 0.5 * kPi;
    P.ldc_c2psi = cos(2. * psi_ldc);
    P.ldc_s2psi = sin(2. * psi_ldc);
  }
  P.n_vec[0] = sin(qS) * cos(phiS);
  P.n_vec[1] = sin(qS) * sin(phiS);
  P.n_vec[2] = cos(qS);
  P.S_vec[0] = sin(qK) * cos(phiK);
  P.S_vec[1] = sin(qK) * sin(phiK);
  P.S_vec[2] = cos(qK);
  P.cosiota_L = cos(lam);
  P.siniota_L = sin(lam);
  P.sin_thK = sin(qK);
  P.cos_thK = cos(qK);
  P.sin_phK = sin(phiK);
  P.cos_phK = cos(phiK);
  P.cos_phS = cos(phiS);
  P.sin_phS = sin(phiS);
  {
    // the per-sample orientation of L (KSParMap.cc:532-574, AAK.cc:425-433) on
    // the basis (1, sin alpha, cos alpha)
    const double cl = P.coslam, sl = P.sinlam;
    const double g_cosqL[3] = {P.cosqK * cl, 0.0, P.sinqK * sl};
    const double g_up[3] = {P.sinqK * P.sinphiK * cl, -P.cosphiK * sl, -P.cosqK * P.sinphiK * sl};
    const double g_dw[3] = {P.sinqK * P.cosphiK * cl, P.sinphiK * sl, -P.cosqK * P.cosphiK * sl};
    const double g_bd[3] = {0.0, (P.cosqK * P.Sdotn - P.cosqS) / P.sinqK * sl, P.sinqS * P.sin_dphiKS * sl};
    for (int k = 0; k < 3; k++) {
      P.g_cosqL[k] = g_cosqL[k];
      P.g_phiLup[k] = g_up[k];
      P.g_phiLdown[k] = g_dw[k];
      P.g_betadown[k] = g_bd[k];
    }
    const double ci = P.cosiota_L, si = P.siniota_L;
    const double Lk[3][3] = {{ci * P.sin_thK * P.cos_phK, ci * P.sin_thK * P.sin_phK, ci * P.cos_thK},
                             {si * P.sin_phK, -si * P.cos_phK, 0.0},
                             {-si * P.cos_thK * P.cos_phK, -si * P.cos_thK * P.sin_phK, si * P.sin_thK}};
    const double *n = P.n_vec, *Sv = P.S_vec;
    auto cross = [](const double *a, const double *b, double *c) {
      c[0] = a[1] * b[2] - a[2] * b[1];
      c[1] = a[2] * b[0] - a[0] * b[2];
      c[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dot = [](const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    double C[3][3], nxS[3];
    for (int k = 0; k < 3; k++) cross(n, Lk[k], C[k]);
    cross(n, Sv, nxS);
    for (int k = 0; k < 3; k++) {
      P.g_rc[k] = dot(C[k], nxS);
      P.g_rs[k] = dot(Lk[k], nxS) - dot(Sv, C[k]);
    }
    P.g_nxL2[0] = dot(C[0], C[0]);
    P.g_nxL2[1] = dot(C[1], C[1]);
    P.g_nxL2[2] = dot(C[2], C[2]);
    P.g_nxL2[3] = 2. * dot(C[0], C[1]);
    P.g_nxL2[4] = 2. * dot(C[0], C[2]);
    P.g_nxL2[5] = 2. * dot(C[1], C[2]);
    P.nxS2 = dot(nxS, nxS);
  }
  P.zeta = par[0] / sp.D / kGpc;
  P.Mphys = sp.M;
  // Bessel evaluation of the mode sum: series for e below kSeriesMaxE.  Only
  // the knots that shape e on the signal (samples below ip + ib: the knots up
  // to the first one past the last signal sample) bound it; past a plunge the
  // AK evolution can run far off before it is cut, and those knots set no
  // sample the mode sum evaluates.  Samples above the bound (spline
  // overshoot) take the recurrence instead (P.e_series).
  const int n_sig = std::min(N, ip + ib);
  const double t_sig = (double)(n_sig - 1 - P.i0) * P.dt;
  double e_hi = 0.0;
  for (int k = 0; k < ks.K; k++) {
    e_hi = std::max(e_hi, ks.ky(0)[k]);
    if (ks.kt()[k] > t_sig) break;
  }
  const double e_bound = e_hi * 1.01 + 0.01;
  P.bessel_series = (nmodes <= kMaxModes && e_bound <= kSeriesMaxE) ? 1 : 0;
  P.e_series = P.bessel_series ? e_bound : 0.0;
  P.mixed = cfg.mixed ? 1 : 0;
  P.xshift = cfg.exp_shift >= 0 ? cfg.exp_shift : (cfg.mixed ? 2 : 0);
  P.full_dense = 1;
  {
    // mixed form: E1 = sin(lam) S - cos(lam) (cos a e1 + sin a e2) and
    // E2 = L x E1 = sin a e1 - cos a e2, with L's precession basis
    // e1 = (z - S cos theta_K) / sin theta_K, e2 = S x e1
    GeomF32 &F = P.gf;
    const double e1[3] = {-P.cos_thK * P.cos_phK, -P.cos_thK * P.sin_phK, P.sin_thK};
    const double e2[3] = {P.sin_phK, -P.cos_phK, 0.0};
    const double pv[3] = {cos(qS) * cos(phiS), cos(qS) * sin(phiS), -sin(qS)};
    const double *n = P.n_vec;
    const double qv[3] = {pv[1] * n[2] - pv[2] * n[1], pv[2] * n[0] - pv[0] * n[2], pv[0] * n[1] - pv[1] * n[0]};
    for (int k = 0; k < 3; k++) {
      F.e1c[k] = (float)(P.siniota_L * P.S_vec[k]);
      F.e1s[k] = (float)(-P.cosiota_L * e2[k]);
      F.e1k[k] = (float)(-P.cosiota_L * e1[k]);
      F.e2s[k] = (float)e1[k];
      F.e2k[k] = (float)(-e2[k]);
      F.n[k] = (float)n[k];
      F.p[k] = (float)pv[k];
      F.q[k] = (float)qv[k];
    }
    F.cos_phS = (float)P.cos_phS;
    F.sin_phS = (float)P.sin_phS;
  }
  if (P.bessel_series) bessel_series_terms(e_bound, nmodes, P.bterms);
  return 0;
}

void bessel_series_table(double *c) {
  for (int k = 0; k <= kBesselKMax; k++) {
    long double t = 1.0L;
    for (int m = 0; m < kBesselMMax; m++) {
      if (m > 0) t /= (long double)m * (long double)(k + m);
      c[k * kBesselMMax + m] = (double)t;
    }
  }
}

void bessel_log_factorials(double *lgf) {
  for (int k = 0; k <= kBesselKMax; k++) lgf[k] = (double)lgammal((long double)k + 1.0L);
}

// Smallest series length whose first omitted term, past the largest term,
// is below 1e-18 of the leading one, for order k at argument x.
static int series_length(int k, double x) {
  const double q = 0.25 * x * x;
  double t = 1.0;
  for (int m = 1; m < kBesselMMax; m++) {
    const double tn = t * q / ((double)m * (double)(k + m));
    if (tn < t && tn < 1e-18) return m;
    t = tn;
  }
  return kBesselMMax;
}

void bessel_series_terms(double e_bound, int n_modes, unsigned char *terms) {
  int m0 = 2;
  for (int k = 0; k <= 4; k++) m0 = std::max(m0, series_length(k, 2.0 * e_bound));
  terms[0] = (unsigned char)m0;
  for (int n = 1; n <= n_modes && n <= kMaxModes; n++) {
    const double x = n * e_bound;
    const int m = std::max(series_length(n + 1, x), series_length(n + 2, x));
    terms[n] = (unsigned char)std::max(2, m);
  }
}

}  // namespace aak
