// aak_fft.hpp -- fused real FFT + noise-weighted inner products (aak_fft.hip).
#pragma once

#include <hip/hip_runtime.h>

namespace aak {

// N = 2M real samples per channel, M = R C (powers of two).  Tables: W_R^e,
// W_C^e (e < R, C), W_M^e = twMhi[e >> 10] twMlo[e & 1023] for e < M and
// W_N^k = twNhi[k >> 10] twNlo[k & 1023] for k <= M, twNr[k2] = W_N^(R k2) for
// k2 < C.  Dp, wp: weighted data and
// weights of each channel in the order of the second pass, position
// k1 * C + k2 for bin k = k1 + R k2 (k < M), position M for the Nyquist bin.
struct FftArgs {
  int N, M, R, C;
  const double2 *twR, *twC, *twMlo, *twMhi, *twNlo, *twNhi, *twNr;
  const double2 *Dp[2];
  const double *wp[2];
};

// Lengths handled (powers of two, 2^8 <= N <= 2^23) and their R, C.
bool fft_supported(int N, int *R, int *C);
// Partial-sum pairs written by launch_fft_pass2.
int fft_partials(const FftArgs &F);
// Twiddle tables of F's lengths: their size (double2 entries), their values
// (host memory, extended-precision roots) and the table pointers of F bound
// to a device copy.
size_t fft_table_size(const FftArgs &F);
void fft_fill_tables(const FftArgs &F, double2 *tw);
void fft_bind_tables(FftArgs *F, const double2 *tw);
// Channels h0 and h1 (16-byte aligned).  Pass 1 writes the 2M-complex scratch Y;
// pass 2 writes fft_partials(F) pairs (sum of <h w, d w>, sum of |h w|^2) to
// partials and, if hw_out is not null, the weighted spectra h~ w ([2][M+1],
// natural order).
void launch_fft_pass1(const FftArgs &F, const double *h0, const double *h1, double2 *Y, hipStream_t s);
void launch_fft_pass2(const FftArgs &F, const double2 *Y, double *partials, double2 *hw_out, hipStream_t s);

// Pass 2 with the final reduction folded in: each row workgroup publishes its
// pair in `partials` (words that read kFinalEmpty until written); one extra
// workgroup, dispatched after all of them, collects the pairs and sums them
// in the order of the separate final reduction (k_dot_final: 256 lanes, lane t adds pairs t, t + 256, ...
// in turn, then a halving tree), writes out[0..1] = 4 x the sums (NaN when
// flags[3] is set), zeroes clear[0..7] (flags and clear may be null) and
// returns every word to kFinalEmpty.  A pair that never arrives within the
// spin cap reads as NaN.  launch_fft_final_reset brings a new buffer to that
// state.
struct FftFinal {
  double *out;
  const int *flags;
  int *clear;
};
constexpr long long kFinalEmpty = 0x7FF4A5A5A5A5A5B1LL;  // a signalling NaN: no arithmetic result
void launch_fft_pass2_final(const FftArgs &F, const double2 *Y, double *partials, const FftFinal &fin, hipStream_t s);
void launch_fft_final_reset(double *partials, int npairs, hipStream_t s);

}  // namespace aak
