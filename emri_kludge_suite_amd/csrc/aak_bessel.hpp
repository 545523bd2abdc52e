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

constexpr int kBesselKMax = 140;  // orders 0..140: harmonics up to kMaxModes plus Taylor orders
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

// ---------------------------------------------------------- Taylor in e
// alpha_n(e) = -n (J_{n-2} - 2e J_{n-1} + 2/n J_n + 2e J_{n+1} - J_{n+2}),
// beta_n(e)  = -n sqrt(1-e^2) (J_{n-2} - 2 J_n + J_{n+2}),  gamma_n(e) = 2 J_n
// (all J at argument n e; J_{-1} = -J_1) are the e-dependent factors of the
// harmonic coefficients a_n, b_n, c_n (AAK.cc:493-504).  Their Taylor
// coefficients in delta = e - e_ref follow from
//   d^j/de^j J_k(ne) = (n/2)^j sum_i (-1)^i C(j,i) J_{k-j+2i}(n e_ref).
constexpr int kTaylorMax = 8;
constexpr int kTaylorOrders = 5 + 2 * kTaylorMax;

// J at orders n-2-p .. n+2+p (out[q * S] = J_{n-2-p+q}) at x = n e_ref, negative
// orders by J_{-m} = (-1)^m J_m.  Runs that reach order 0 (low harmonics,
// where J_0 ~ 1) use Miller's recurrence normalised with J_0 + 2 sum J_2k = 1;
// the others take the two highest orders from the series and recur down
// (their values are small enough that the prefactor's relative error does not
// matter absolutely).  M: series length valid for orders >= n+1.
template <int S = 1>
AAK_HD void bessel_run(int n, int p, double e_ref, const double *ctab, const double *lgf, int M, double *out) {
  const int kmin = n - 2 - p, kmax = n + 2 + p, cnt = kmax - kmin + 1;
  const double x = n * e_ref;
  if (!(e_ref > 1e-150)) {
    for (int q = 0; q < cnt; q++) out[q * S] = (kmin + q == 0) ? 1.0 : 0.0;
    return;
  }
  if (kmin <= 0) {
    int top = kmax + 30 + (int)x;
    top += (top & 1);
    const double inv = 1.0 / x;
    double jp1 = 0.0, j = 1.0e-280, sum = 0.0;
    for (int k = top; k >= 1; k--) {
      if (k <= kmax) out[(k - kmin) * S] = j;
      if ((k & 1) == 0) sum += 2.0 * j;
      const double jm1 = (2.0 * k) * inv * j - jp1;
      jp1 = j;
      j = jm1;
      if (fabs(j) > 1.0e250) {
        j *= 1.0e-250;
        jp1 *= 1.0e-250;
        sum *= 1.0e-250;
        for (int q = k - kmin; q < cnt && q >= 0; q++) out[q * S] *= 1.0e-250;
      }
    }
    out[(-kmin) * S] = j;
    sum += j;
    const double norm = 1.0 / sum;
    for (int q = -kmin; q < cnt; q++) out[q * S] *= norm;
    for (int k = kmin; k < 0; k++) {
      const int m = -k;
      const double v = out[(m - kmin) * S];
      out[(k - kmin) * S] = (m & 1) ? -v : v;
    }
    return;
  }
  const double y = -0.25 * x * x, lhx = log(0.5 * x);
  const double ptop = exp(kmax * lhx - lgf[kmax]);
  const double jt = ptop * bessel_series_sum(ctab + kmax * kBesselMMax, M, y);
  const double jt1 = ptop * ((double)kmax / (0.5 * x)) * bessel_series_sum(ctab + (kmax - 1) * kBesselMMax, M, y);
  out[(cnt - 1) * S] = jt;
  out[(cnt - 2) * S] = jt1;
  const double inv = 1.0 / x;
  double jk1 = jt, jk = jt1;
  const int klo = kmin > 0 ? kmin : 0;
  for (int k = kmax - 1; k > klo; k--) {
    const double jm = (2.0 * k) * inv * jk - jk1;
    out[(k - 1 - kmin) * S] = jm;
    jk1 = jk;
    jk = jm;
  }
}

// Taylor coefficients 0..p of sqrt(1 - (e_ref + d)^2) into sq.
AAK_HD void taylor_sqrt1me2(double e_ref, int p, double *sq) {
  const double g0 = 1. - e_ref * e_ref, g1 = -2. * e_ref, g2 = -1.;
  sq[0] = sqrt(g0);
  for (int j = 1; j <= p; j++) {
    double acc = (j == 1 ? g1 : (j == 2 ? g2 : 0.0));
    for (int q = 1; q < j; q++) acc -= sq[q] * sq[j - q];
    sq[j] = acc / (2. * sq[0]);
  }
}

// Taylor coefficients of alpha_n, beta_n, gamma_n at e_ref to order PO into
// coef[f * (PO + 1) + j] (f = alpha, beta, gamma).  Jr[q * S] = J_{n-2-PO+q}(n e_ref)
// for q < 5 + 2 PO; sq: Taylor coefficients of sqrt(1 - e^2) at e_ref.  The
// signed binomial sums of the derivative formula are formed as iterated
// differences, D_j(k) = D_{j-1}(k-1) - D_{j-1}(k+1) with D_0 = J, on one row that
// shrinks by two orders per step.
template <int PO, int S, class Out>
AAK_HD void taylor_amplitudes_p(int n, double e_ref, const double *sq, const double *Jr, Out out) {
  constexpr int L = 5 + 2 * PO;
  double row[L];
  for (int q = 0; q < L; q++) row[q] = Jr[q * S];
  // coefficients are emitted order by order as their differences appear:
  // t_k = (n/2)^j / j! D_j(n - 2 + k) of the current order, u[j] = t_0 - 2 t_2 + t_4
  // of every order so far (the beta expansion convolves them with sq)
  double u[PO + 1];
  double t1_prev = 0.0, t3_prev = 0.0;
  double hn = 1.0;
  const double dn = (double)n, two_n = 2. / dn;
  for (int j = 0; j <= PO; j++) {
    const double t0 = hn * row[PO - j], t1 = hn * row[1 + PO - j], t2 = hn * row[2 + PO - j];
    const double t3 = hn * row[3 + PO - j], t4 = hn * row[4 + PO - j];
    const double e1 = e_ref * t1 + (j ? t1_prev : 0.0);
    const double e3 = e_ref * t3 + (j ? t3_prev : 0.0);
    out(0, j, -dn * (t0 - 2. * e1 + two_n * t2 + 2. * e3 - t4));
    u[j] = t0 - 2. * t2 + t4;
    double bsum = 0.0;
    for (int q = 0; q <= j; q++) bsum += sq[q] * u[j - q];
    out(1, j, -dn * bsum);
    out(2, j, 2. * t2);
    t1_prev = t1;
    t3_prev = t3;
    for (int q = 0; q + 2 < L - 2 * j; q++) row[q] = row[q] - row[q + 2];
    hn = hn * (0.5 * n) * (1.0 / (j + 1));
  }
}

// Store of taylor_amplitudes_p into coef[f * (PO + 1) + j].
template <int PO, class T>
struct CoefStore {
  T *coef;
  AAK_HD void operator()(int f, int j, double v) const { coef[f * (PO + 1) + j] = (T)v; }
};

// Runtime-order form of taylor_amplitudes_p (coef holds 3 (p + 1) values).
AAK_HD void taylor_amplitudes(int n, int p, double e_ref, const double *sq, const double *ctab, const double *lgf,
                              int M, double *coef) {
  double Jr[kTaylorOrders];
  bessel_run(n, p, e_ref, ctab, lgf, M, Jr);
  switch (p) {
    case 0: taylor_amplitudes_p<0, 1>(n, e_ref, sq, Jr, CoefStore<0, double>{coef}); break;
    case 1: taylor_amplitudes_p<1, 1>(n, e_ref, sq, Jr, CoefStore<1, double>{coef}); break;
    case 2: taylor_amplitudes_p<2, 1>(n, e_ref, sq, Jr, CoefStore<2, double>{coef}); break;
    case 3: taylor_amplitudes_p<3, 1>(n, e_ref, sq, Jr, CoefStore<3, double>{coef}); break;
    case 4: taylor_amplitudes_p<4, 1>(n, e_ref, sq, Jr, CoefStore<4, double>{coef}); break;
    case 5: taylor_amplitudes_p<5, 1>(n, e_ref, sq, Jr, CoefStore<5, double>{coef}); break;
    case 6: taylor_amplitudes_p<6, 1>(n, e_ref, sq, Jr, CoefStore<6, double>{coef}); break;
    case 7: taylor_amplitudes_p<7, 1>(n, e_ref, sq, Jr, CoefStore<7, double>{coef}); break;
    default: taylor_amplitudes_p<8, 1>(n, e_ref, sq, Jr, CoefStore<8, double>{coef}); break;
  }
}

// Horner evaluation of the three expansions (layout of taylor_amplitudes) at delta.
AAK_HD void taylor_eval(const double *coef, int p, double d, double *abg) {
  const int s = p + 1;
  double al = coef[p], be = coef[s + p], ga = coef[2 * s + p];
  for (int j = p - 1; j >= 0; j--) {
    al = fma(al, d, coef[j]);
    be = fma(be, d, coef[s + j]);
    ga = fma(ga, d, coef[2 * s + j]);
  }
  abg[0] = al;
  abg[1] = be;
  abg[2] = ga;
}

// Remainder bounds of the expansion: FP64 sums, and FP32 sums (whose
// coefficients and arithmetic carry relative errors near 6e-8 anyway).
constexpr double kTaylorTolFp64 = 1e-17;
constexpr double kTaylorTolFp32 = 1e-9;

// Taylor order for harmonics up to n_max and |delta| <= del with remainder
// bound at most tol (-1: none <= kTaylorMax).
AAK_HD int taylor_order(int n_max, double del, double tol = kTaylorTolFp64) {
  const double nm = (double)n_max, z = nm * del;
  double term = 7.0 * nm * z;  // 7 n (n delta)^{p+1} / (p+1)! at p = 0
  for (int q = 0; q <= kTaylorMax; q++) {
    if (term <= tol) return q;
    term = term * z / (q + 2);
  }
  return -1;
}

}  // namespace aak
