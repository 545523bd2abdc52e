// Host side of the command-line waveform generator: the settings/parameters
// file and the analytic eLISA noise curve of its SNR.
//
//   aak_load_setpar   KSTools.cc:26-114 (LoadSetPar): one path, five booleans
//                     and nineteen numbers, one entry per line; lines holding
//                     '#' and empty lines are skipped; only the first token of
//                     a line is read.  Same messages, same entry numbering,
//                     same derived dt (T in 365-day years) and p (from the
//                     reference GW frequency f = Omega_phi / pi).
//   aak_elisa_noise   KSTools.cc:116-125, the eLISA sensitivity of
//                     Amaro-Seoane et al. (2013), arm length 1e9 m.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/gpuAAK.h"
#include "aak_math.hpp"

namespace aak {
int set_last_error(int code, const std::string &msg);
}

namespace {
// LoadSetPar reports failure as 0 with a message
int parse_fail(const std::string &msg) {
  aak::set_last_error(0, msg);
  return 0;
}
}  // namespace

extern "C" {

int aak_load_setpar(const char *filename, aak_setpar *out) {
  if (!filename || !out) {
    return parse_fail("null argument");
  }
  std::string path;
  bool flags[5] = {false, false, false, false, false};
  double num[19];
  int n_str = 0, n_bool = 0, n_num = 0;
  std::ifstream file(filename);
  std::string line;
  char msg[160];
  while (std::getline(file, line)) {
    if (line.empty() || line.find('#') != std::string::npos) continue;
    std::istringstream in(line);
    if (n_str < 1) {
      in >> path;
      if (in.fail()) {
        snprintf(msg, sizeof msg, "Parsing error: Check entry %d in settings/parameters file", 1);
        return parse_fail(msg);
      }
      n_str++;
    } else if (n_bool < 5) {
      in >> std::boolalpha >> flags[n_bool];
      if (in.fail()) {
        snprintf(msg, sizeof msg, "Parsing error: Check entry %d in settings/parameters file", n_bool + 2);
        return parse_fail(msg);
      }
      n_bool++;
    } else if (n_num < 19) {
      in >> num[n_num];
      if (in.fail()) {
        snprintf(msg, sizeof msg, "Parsing error: Check entry %d in settings/parameters file", n_num + 6);
        return parse_fail(msg);
      }
      n_num++;
    } else {
      return parse_fail("Parsing error: Too many entries in settings/parameters file");
    }
  }
  if (n_str + n_bool + n_num < 25) {
    return parse_fail("Parsing error: Too few entries in settings/parameters file");
  }
  if (path.size() >= sizeof(out->path)) {
    return parse_fail("Parsing error: path in settings/parameters file is too long");
  }
  std::memset(out, 0, sizeof(*out));
  std::memcpy(out->path, path.c_str(), path.size() + 1);
  out->backint = flags[0];
  out->LISA = flags[1];
  out->traj = flags[2];
  out->SNR = flags[3];
  out->timing = flags[4];
  out->length = (int)num[0];
  double *dst[18] = {&out->dt, &out->p, &out->T, &out->f, &out->T_fit, &out->mu, &out->M, &out->s, &out->e,
                     &out->iota, &out->gamma, &out->psi, &out->theta_S, &out->phi_S, &out->theta_K,
                     &out->phi_K, &out->alpha, &out->D};
  for (int k = 0; k < 18; k++) *dst[k] = num[k + 1];
  // unspecified step: duration T in years of 365 days over the sample count
  if (out->dt < 0) out->dt = out->T * 31536000. / out->length;
  // unspecified p: Keplerian p of the reference GW frequency, f = Omega_phi / pi
  if (out->p < 0) out->p = (1. - out->e * out->e) / pow(M_PI * out->M * aak::kMsunSec * out->f, 2. / 3.);
  return 1;
}

double aak_elisa_noise(double f) {
  const double L = 1.e9;                                   // arm length (m)
  const double acc = 1.37e-32 * (1. + 1.e-4 / f) / pow(f, 4.);  // acceleration noise
  const double sn = 5.25e-23, omn = 6.28e-23;              // shot and other optical-path noise
  const double x = 2. * L * f / 0.41 / aak::kC;
  return 20. / 3. * (4. * acc + sn + omn) / L / L * (1. + x * x);
}

}  // extern "C"
