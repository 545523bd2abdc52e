// AAK_Waveform: the suite's command-line waveform generator on the GPU.
//
//   AAK_Waveform <settings/parameters file>
//
// Same file format and outputs as the reference's executable
// (src/exec/AAK_Waveform.cc:14-134):
//   <path>_wave.dat   t, h_I, h_II per sample ("%14.12e %14.12e %14.12e")
//   <path>_traj.dat   t, p/M, e, iota, E, L_z, Q up to the plunge, every
//                     i_RR-th sample (i_RR = 0.001 M^2/mu in steps of dt) and
//                     the last i_RR samples; p recovered from the mapped
//                     azimuthal frequency, (1 - e^2) (M Omega_phi)^(-2/3)
//   <path>_info.txt   SNR_I, SNR_II, SNR under the analytic eLISA noise curve
//                     (bins 1 .. (length+1)/2 - 1 with 1e-5 < f < 3 Hz) and
//                     the generation's clock ticks and seconds (process CPU
//                     time, as clock() counts it)
// The waveform comes from aak_gpu_gen_AAK; the SNRs are the likelihood
// stage's <h|h> on the GPU, one channel at a time (the other weighted by 0).
#include <cmath>
#include <cstdio>
#include <ctime>
#include <string>
#include <vector>

#include "../../include/gpuAAK.h"

namespace {

constexpr double kF_min = 1.e-5, kF_max = 3.0;  // KSTools.h:18-19
// G Msun / c^3 (Constants.h:4-10)
const double kMsunSec = 6.67259e-11 * 1.98892e30 / (299792458.0 * 299792458.0 * 299792458.0);

bool readable(const std::string &name) {
  FILE *f = fopen(name.c_str(), "r");
  if (!f) return false;
  fclose(f);
  return true;
}

FILE *open_output(const std::string &name) {
  if (readable(name)) fprintf(stderr, "Output warning: Overwriting %s\n", name.c_str());
  FILE *f = fopen(name.c_str(), "w");
  if (!f) fprintf(stderr, "Output error: Cannot write %s\n", name.c_str());
  return f;
}

int gpu_error(const char *what) {
  fprintf(stderr, "AAK_Waveform: %s failed: %s\n", what, aak_last_error());
  return 1;
}

// 4 <h|h> of one channel: data zero, the other channel weighted by zero
int channel_power(aak_handle h, const aak_setpar &s, const std::vector<double> &w, int channel, double *hh) {
  const int nb = s.length / 2 + 1;
  std::vector<double> zero_c(2 * (size_t)nb, 0.0), zero_w((size_t)nb, 0.0);
  const double *w1 = channel == 0 ? w.data() : zero_w.data();
  const double *w2 = channel == 0 ? zero_w.data() : w.data();
  if (aak_input_data(h, zero_c.data(), zero_c.data(), w1, w2, nb)) return gpu_error("input_data");
  if (aak_gpu_gen_AAK(h, s.iota, s.s, s.p, s.e, s.M, s.mu, s.gamma, s.psi, s.alpha, s.theta_S, s.phi_S,
                      s.theta_K, s.phi_K, s.D))
    return gpu_error("gpu_gen_AAK");
  double like[2];
  if (aak_likelihood(h, like)) return gpu_error("Likelihood");
  *hh = like[1];
  return 0;
}

}  // namespace

int main(int argc, char *argv[]) {
  if (argc < 2) {
    fprintf(stderr, "Input error: Specify path to settings/parameters file\n");
    return 0;
  }
  if (!readable(argv[1])) {
    fprintf(stderr, "Input error: Cannot read settings/parameters file\n");
    return 0;
  }
  aak_setpar s;
  if (!aak_load_setpar(argv[1], &s)) {
    fprintf(stderr, "%s\n", aak_last_error());
    return 0;
  }
  if (s.length < 2) {
    fprintf(stderr, "Input error: length must be at least 2\n");
    return 0;
  }
  const size_t n = (size_t)s.length;
  aak_handle h = nullptr;
  if (aak_create(s.T_fit, s.length, s.length, s.dt, s.dt, s.LISA, s.backint, &h)) return gpu_error("create");

  // ----- waveform -----
  clock_t ticks = clock();
  std::vector<double> t(n + 2), hI(n + 2), hII(n + 2);
  if (aak_gpu_gen_AAK(h, s.iota, s.s, s.p, s.e, s.M, s.mu, s.gamma, s.psi, s.alpha, s.theta_S, s.phi_S,
                      s.theta_K, s.phi_K, s.D))
    return gpu_error("gpu_gen_AAK");
  if (aak_get_waveform(h, t.data(), hI.data(), hII.data())) return gpu_error("GetWaveform");
  ticks = clock() - ticks;
  const double secs = (double)ticks / CLOCKS_PER_SEC;

  const std::string base(s.path);
  FILE *f = open_output(base + "_wave.dat");
  if (!f) return 1;
  for (size_t i = 0; i < n; i++) fprintf(f, "%14.12e %14.12e %14.12e\n", t[i], hI[i], hII[i]);
  fclose(f);

  // ----- trajectory -----
  if (s.traj) {
    std::vector<double> tt(n), pp(n), ee(n);
    const int rows = aak_get_inspiral(h, tt.data(), pp.data(), ee.data(), s.length);
    if (rows < 0) return gpu_error("inspiral");
    f = open_output(base + "_traj.dat");
    if (!f) return 1;
    const int i_rr = (int)(0.001 * (kMsunSec * s.M * s.M / s.mu) / s.dt) + 1;
    int i_max = 0;
    while (i_max < rows && pp[i_max] > 0.) i_max++;
    for (int i = 0; i < i_max; i++) {
      if (i % i_rr == 0 || i + i_rr >= i_max) {
        double elq[3];
        if (aak_kerr_constants(pp[i], ee[i], s.iota, s.s, elq)) return gpu_error("Kerr constants");
        fprintf(f, "%14.12e %14.12e %14.12e %14.12e %14.12e %14.12e %14.12e\n", tt[i], pp[i], ee[i], s.iota, elq[0],
                elq[1], elq[2]);
      }
    }
    fclose(f);
  }

  // ----- SNR and timing -----
  if (s.SNR || s.timing) {
    double hI2 = 0., hII2 = 0.;
    if (s.SNR) {
      // w_k^2 = dt / (length S_n(f_k)): 4 |H_k w_k|^2 summed is the
      // reference's InnProd of the FFTW half-complex spectra
      std::vector<double> w(n / 2 + 1, 0.0);
      for (int k = 1; k < (s.length + 1) / 2; k++) {
        const double fk = k / s.dt / s.length;
        if (fk > kF_min && fk < kF_max) w[k] = sqrt(s.dt / s.length / aak_elisa_noise(fk));
      }
      if (channel_power(h, s, w, 0, &hI2) || channel_power(h, s, w, 1, &hII2)) return 1;
    }
    f = open_output(base + "_info.txt");
    if (!f) return 1;
    if (s.SNR) fprintf(f, "SNR_I: %.1f\nSNR_II: %.1f\nSNR: %.1f\n\n", sqrt(hI2), sqrt(hII2), sqrt(hI2 + hII2));
    if (s.timing) fprintf(f, "Clock ticks: %.2e\nSeconds: %.2fs\n\n", (double)ticks, secs);
    fclose(f);
  }
  aak_destroy(h);
  return 0;
}
