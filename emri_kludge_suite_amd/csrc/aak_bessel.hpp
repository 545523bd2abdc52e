// aak_bessel.hpp -- J_{n-2..n+2}(n e) for the AAK harmonic sum.
//
// The reference evaluates five Bessel functions per (sample, harmonic) with
// GSL's gsl_sf_bessel_Jn_array (AAK.cc:495-501), i.e. one backward recurrence
// from far above n down to order 0 per harmonic: O(n) work for harmonic n and
// O(n_max^2) per sample.  Here the two highest orders come from the power
// series
//     J_k(x) = (x/2)^k / k! * sum_m c_m(k) y^m,   y = -x^2/4,
//     c_m(k) = 1 / (m! (k+1)(k+2)...(k+m)),
// evaluated by Horner's rule with tabulated coefficients, and the three lower
// orders by the backward recurrence J_{k-1} = (2k/x) J_k - J_{k+1}, which is
// stable downwards for k > x.  The number of series terms per harmonic is
// fixed for a whole waveform from an upper bound on e, so every lane of a
// wave runs the same loop.  For e <= kSeriesMaxE the series has no harmful
// cancellation (I_k(x) <= J_k(x)-scale), and the relative error of each J is a
// few ulp; above it the caller uses the Miller recurrence instead.
#pragma once

#include <cmath>

#include "aak_math.hpp"

namespace aak {

constexpr int kBesselKMax = 130;  // orders 0..130: harmonics up to kMaxModes
constexpr int kBesselMMax = 64;   // series terms 0..63
constexpr int kMaxModes = 128;
constexpr double kSeriesMaxE = 0.66;

// c_m(k) for k = 0..kBesselKMax, m = 0..kBesselMMax-1, row-major by k.
void bessel_series_table(double *c);
// log(k!) for k = 0..kBesselKMax.
void bessel_log_factorials(double *lgf);

// Series length per harmonic for eccentricities up to e_bound (relative
// truncation below 1e-17): terms[n] for n = 1..n_modes, terms[0] is the
// length used for the five low orders of harmonics 1 and 2.
void bessel_series_terms(double e_bound, int n_modes, unsigned char *terms);

// Horner sum_{m<M} c_m(k) y^m.
AAK_HD double bessel_series_sum(const double *ck, int M, double y) {
  double s = ck[M - 1];
  for (int m = M - 2; m >= 0; m--) s = fma(s, y, ck[m]);
  return s;
}

// J_{n-2..n+2}(n e) into J[0..4]; J_{-1} = -J_1.  lgf[k] = log(k!).
// e must lie in [0, kSeriesMaxE]; M is the series length of harmonic n and
// M0 that of the low orders (n <= 2).
AAK_HD void bessel_five_series(int n, double e, const double *ctab, const double *lgf, int M, int M0,
                                      double *J) {
  const double x = n * e;
  if (!(e > 1e-150)) {
    for (int q = 0; q < 5; q++) J[q] = (n - 2 + q == 0) ? 1.0 : 0.0;
    return;
  }
  const double y = -0.25 * x * x;
  const double lhx = log(0.5 * x);
  if (n <= 2) {
    // orders max(0, n-2) .. n+2 straight from the series
    for (int q = 0; q < 5; q++) {
      const int k = n - 2 + q;
      if (k < 0) continue;
      const double pref = (k == 0) ? 1.0 : exp(k * lhx - lgf[k]);
      J[q] = pref * bessel_series_sum(ctab + k * kBesselMMax, M0, y);
    }
    if (n == 1) J[0] = -J[2];
    return;
  }
  const int k1 = n + 1, k2 = n + 2;
  const double p1 = exp(k1 * lhx - lgf[k1]);
  const double p2 = p1 * (0.5 * x) / k2;
  const double j1 = p1 * bessel_series_sum(ctab + k1 * kBesselMMax, M, y);
  const double j2 = p2 * bessel_series_sum(ctab + k2 * kBesselMMax, M, y);
  const double inv = 1.0 / x;
  const double j0 = (2.0 * k1) * inv * j1 - j2;        // J_n
  const double jm1 = (2.0 * n) * inv * j0 - j1;        // J_{n-1}
  const double jm2 = (2.0 * (n - 1)) * inv * jm1 - j0; // J_{n-2}
  J[0] = jm2;
  J[1] = jm1;
  J[2] = j0;
  J[3] = j1;
  J[4] = j2;
}

}  // namespace aak
