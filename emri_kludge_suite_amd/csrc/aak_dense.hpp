// aak_dense.hpp -- host/device contract of the dense (per-sample) stage.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "aak_bessel.hpp"

namespace aak {

// Channel stride of the time-domain buffer: N+2 samples rounded up to 128
// bytes, so that both channels start on a cache line.
AAK_HD size_t h_stride(int N) { return ((size_t)N + 2 + 15) & ~(size_t)15; }

// Device knot table: y of e, v, M, S (4K) | their spline coefficients c (4K) |
// head knot times (KnotGrid::steps).
constexpr int kTableArrays = 8;

// Knot times of the device knot table.  Knot k has evolution index
// j = j_start + k and time
//   j <  j0            dt_large * (j - j0)                (backward evolution)
//   j0 <= j < j0+steps dt_map[j - j0]                     (stored "head")
//   j >= j0+steps      dt_map[steps-1] + dt_large * (j - j0 - steps + 1)
// exactly as the host evolution computes them (AAK.cc:114-121, 224-231), so
// host and device agree bit for bit without shipping K times.
constexpr int kRatesBlock = 256;  // samples per k_rates workgroup

struct KnotGrid {
  int K;           // knots in the table
  int head_off;    // knot index of dt_map[0]  (j0 - j_start)
  int steps;       // head length
  int flat;        // 1: the evolution's knots beyond K-1 bound flat intervals (value y[K-1])
  double dt_large, t_head_last;  // spacing, dt_map[steps-1]
};

AAK_HD double knot_time(const KnotGrid &g, const double *head, int k) {
  const int jr = k - g.head_off;
  if (jr < 0) return g.dt_large * (double)jr;
  if (jr < g.steps) return head[jr];
  return g.t_head_last + g.dt_large * (double)(jr - g.steps + 1);
}

// Largest k in [0, K-2] with knot_time(k) <= t (0 if t precedes every knot):
// the interval a bisection over the knot times would return.
AAK_HD int knot_index(const KnotGrid &g, const double *head, double t) {
  int k;
  if (t < 0.0) {
    k = g.head_off + (int)floor(t / g.dt_large);
  } else if (t < g.t_head_last) {
    int lo = 0, hi = g.steps - 1;
    while (hi > lo + 1) {
      const int m = (hi + lo) >> 1;
      if (head[m] > t) hi = m;
      else lo = m;
    }
    k = g.head_off + lo;
  } else {
    k = g.head_off + g.steps - 1 + (int)floor((t - g.t_head_last) / g.dt_large);
  }
  if (k > g.K - 2) k = g.K - 2;
  if (k < 0) k = 0;
  while (k < g.K - 2 && knot_time(g, head, k + 1) <= t) k++;
  while (k > 0 && knot_time(g, head, k) > t) k--;
  return k;
}

// Cubic of one spline interval from the knot values and second-derivative
// coefficients, in the evaluation order of the reference's spline.
AAK_HD void spline_interval(double x0, double x1, double y0, double y1, double c0, double c1, double *b,
                                   double *d) {
  const double h = x1 - x0;
  const double dy = y1 - y0;
  *b = (dy / h) - h * (c1 + 2.0 * c0) / 3.0;
  *d = (c1 - c0) / (3.0 * h);
}

// Single-precision copies of the response constants, for the mixed form's
// per-sample geometry.
// Constants of the mixed form's per-sample response, built from unit
// vectors: the orbital-plane pair E1 (the spin's projection, from which the
// pericentre angle gamma is measured) and E2 = L x E1, both affine in
// (sin alpha, cos alpha): E1 = e1c + e1s sin a + e1k cos a, E2 = e2s sin a +
// e2k cos a; the source direction n; the sky polarisation basis (p, q) of
// the LISA = false response.
struct GeomF32 {
  float e1c[3], e1s[3], e1k[3], e2s[3], e2k[3], n[3], p[3], q[3];
  float cos_phS, sin_phS;
};

// Per-waveform scalars of the dense stage, filled on the host.
struct DenseParams {
  int N;           // samples (vlength)
  int i0;          // index of t = 0 on the dense grid
  int K;           // spline knots
  int ip, ib;      // plunge index and taper length in samples
  int last_fold;   // last forward step index folded (min(N-1, ip+ib))
  int guard_last;  // last sample of the Phidot guard (last_fold, or N-1 for phase output)
  int nmodes;
  int mich;        // LISA response (1) or polarisation rotation (0)
  double dt;
  double coslam;   // cos(iota)
  double x0[3];    // gamma0, Phi0, alpha0
  double t_first;  // t[0]  (first knot)
  double t_last;   // t[N]  (last knot)
  // mode-sum constants (computed on the host with the reference's expressions)
  double lam, sinlam, qS, cosqS, sinqS, qK, cosqK, sinqK, phiS, phiK, cosphiK, sinphiK;
  double halfsqrt3, Sdotn, sin_dphiKS, cos_dphiKS;
  double psi_c0, psi_c1, psi_c2, psi_c3;  // constant parts of the polarisation angle (LISA)
  double ldc_c2psi, ldc_s2psi;    // response without LISA
  double n_vec[3], S_vec[3];      // sky and spin unit vectors
  double cosiota_L, siniota_L;    // cos/sin of the inclination in the L vector
  double sin_thK, cos_thK, sin_phK, cos_phK;
  double cos_phS, sin_phS;        // cos/sin(phi_S)
  // orientation terms of one sample as affine functions of (sin alpha, cos alpha):
  // X = c[0] + c[1] sin(alpha) + c[2] cos(alpha), alpha the Lense-Thirring phase
  double g_cosqL[3], g_phiLup[3], g_phiLdown[3], g_betadown[3];
  double g_rc[3], g_rs[3];        // (n x L).(n x S) and L.(n x S) - S.(n x L)
  double g_nxL2[6];               // |n x L|^2 on 1, sa^2, ca^2, sa, ca, sa ca
  double nxS2;                    // |n x S|^2
  double zeta, Mphys;             // amplitude scale and M in solar masses
  double t_plunge, t_zero;
  KnotGrid grid;
  int mixed;                      // 1: FP32 harmonic loop (mixed precision)
  int xshift;                     // expansion blocks of the mode sum span 256 << xshift samples
  int full_dense;                 // 1: every dense array at every sample (single waveforms: phase
                                  // output, dense arrays); 0: only what a likelihood reads
  GeomF32 gf;                     // response constants in FP32 (mixed form)
  int bessel_series;              // 1: series + recurrence, 0: Miller from above
  double e_series;  // series valid for e <= e_series (0: recurrence only)
  unsigned char bterms[kMaxModes + 1];  // series length per harmonic ([0]: orders 0..4)
};

// Exact-order parallel fold bookkeeping.
constexpr int kGuardGrid = 64;  // workgroups of the Phidot guard
constexpr int kFoldThreads = 256;
constexpr int kFoldPerThread = 16;
constexpr int kFoldBlock = kFoldThreads * kFoldPerThread;
constexpr int kFoldMaxItems = 48;
constexpr int kFoldMaxTies = 32;

struct FoldItem {
  int kind;  // 0 = run of steps on one binade grid, 1 = single exact step
  int k;     // binade exponent (frexp convention) of the run
  int sgn;   // +1 / -1
  int first; // first step (block-local)
  int tie_first;
  int blk;            // block of the item (set when the chain gathers the lists)
  long long R;        // run total of rounded increments (ties rounded down)
  long long m_start;  // integer representation of the exact run start (set by the chain)
  double d;           // increment of a single exact step (kind 1)
};

struct FoldTie {
  int step;  // block-local
  int up;    // resolved by the chain: 1 if the tie rounds away from the floor
  long long P;  // inclusive run prefix at the tie (ties rounded down)
};

constexpr int kLW = 8;  // look-back words per block and phase of the single-pass fold

// What k_guard clears before the folds of a waveform: the look-back words,
// replay marks and tickets of both fold directions.
struct FoldReset {
  long long *words;
  int nwords;
  int *status;    // [2][3]
  unsigned *ctr;  // [2]
  int force;      // 1: every phase takes the serial replay (test hook)
};

// One direction of the three phase folds (kernel argument).
struct FoldArgs {
  const double *rate[3];
  double *out[3];
  double x0[3];
  int i0, dir, nsteps, nblk;
  double dt;
};

struct FoldWork {
  double *bsum, *bmin, *bmax, *xstart;  // [3][nblk]
  double *E;                            // [3]
  int *nitems, *nties;                  // [3][nblk]
  FoldItem *items;                      // [3][nblk][kFoldMaxItems]
  FoldTie *ties;                        // [3][nblk][kFoldMaxTies]
  int *status;                          // [3]: 0 ok, 1 overflow/fallback (one fold direction)
  long long *look;                      // [3][nblk][kLW]: look-back words of the single-pass fold
  unsigned *look_ctr;                   // [1]: block tickets of the single-pass fold
  int *fallbacks;                       // serial replays since the last counter reset
  int force_replay;                     // test hook: 1 every phase takes the serial replay, 2 the middle block fails
};

// Device buffers of one waveform slot.
struct DenseBuffers {
  double *knots;     // device knot table: y (4K) | c (4K) | head times
  double *e, *v, *M, *S;           // [N+1]
  double *rate[3];                 // gimdot, Phidot, alpdot [N+1]
  double *phase[3];                // gim, Phi, alp [N+1]
  double *amp;                     // [N+1] mode-sum amplitude (Omega_phi M Msun)^(2/3) zeta
  double *nu0;                     // [1] radial frequency at sample 0 from the PN maps
  double *h, *h2;                  // time domain hI, hII: N+2 samples each, h2 = h + h_stride(N)
  double *t;                       // [N+2]
  int *guard_work;                 // [kGuardGrid + 1]: last good index of each guard segment (-2: not yet), ticket
  int *flags;                      // [8]: [0] non-positive Phidot count, [1] fold replays, [2] per-sample mode-sum
                                   // blocks, [3] non-positive Phidot at t = 0 (waveform rejected), [4] expansion
                                   // blocks split per 256 samples, [5] 0x7fffffff minus the first non-positive
                                   // Phidot sample (0: none), [6..7] unused
  double *ms_coef;                 // per slot nmodes x 3 x (kTaylorMax+1) doubles; order p uses [nmodes][3][p+1] (float in mixed mode); slots: expansion blocks, then 256-sample blocks of split expansions
  int *ms_p;                       // [256-sample blocks] Taylor order (-1: per-sample path)
  int *ms_slot;                    // [256-sample blocks] coefficient slot of the block's expansion
  int *ms_list;                    // [256-sample blocks] blocks of the per-sample path (count in flags[2])
  int *ms_split;                   // [expansion blocks] blocks expanded again per 256 samples (count in flags[4])
  double *ms_eref;                 // [slots] expansion point
  const double *bessel_c;          // series coefficients c_m(k)
  const double *bessel_lgf;        // log(k!)
};

// Launch wrappers (aak_kernels.hip).
void launch_rates(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
// mode 0: single pass with a look-back (k_fold_resolve); 1: three passes
// (k_fold_local, k_fold_chain, k_fold_local)
// defer_replay (single-pass form only): instead of launching k_fold_replay,
// store the direction's arguments there for launch_modesum to carry out
int fold_exact(const DenseParams &P, const DenseBuffers &B, const FoldWork &W, int nblk_cap, int dir, int nsteps,
               int mode, FoldArgs *defer_replay, hipStream_t s);
void launch_fold_fill(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
void launch_phidot_guard(const DenseParams &P, const DenseBuffers &B, const FoldReset &R, hipStream_t s);
int modesum_blocks(const DenseParams &P);  // mode-sum blocks (256 samples each)
int modesum_slots(const DenseParams &P);   // coefficient slots: expansion blocks + 256-sample blocks
// replay (optional): a fold direction whose serial replay check the setup
// kernel carries out in three extra workgroups (it reads no phase)
void launch_modesum(const DenseParams &P, const DenseBuffers &B, const FoldArgs *replay, const FoldWork *fw,
                    hipStream_t s);
void launch_clear_flags(const DenseBuffers &B, hipStream_t s);
void launch_phase_out(const DenseParams &P, const DenseBuffers &B, double *out, hipStream_t s);
// flags (optional): the waveform's counters; flags[3] != 0 (rejected on the
// device) makes both outputs NaN.
void launch_weighted_dot(const double2 *hf, const double2 *data, const double *w, int nbins, double2 *hw_out,
                         double *partials, double *out, const int *flags, int *clear, hipStream_t s);
// 4 x the sums of nparts (dh, hh) pairs into out[0..1].
void launch_dot_final(const double *partials, int nparts, double *out, const int *flags, int *clear, hipStream_t s);

}  // namespace aak
