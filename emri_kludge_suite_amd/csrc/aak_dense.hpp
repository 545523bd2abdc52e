// aak_dense.hpp -- host/device contract of the dense (per-sample) stage.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace aak {

// Per-waveform scalars of the dense stage, filled on the host.
struct DenseParams {
  int N;           // samples (vlength)
  int i0;          // index of t = 0 on the dense grid
  int K;           // spline knots
  int ip, ib;      // plunge index and taper length in samples
  int last_fold;   // last forward step index folded (min(N-1, ip+ib))
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
  double zeta, Mphys;             // amplitude scale and M in solar masses
  double t_plunge, t_zero;
};

// Exact-order parallel fold bookkeeping.
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
  int pad0;
  long long R;        // run total of rounded increments (ties rounded down)
  long long m_start;  // integer representation of the exact run start (set by the chain)
};

struct FoldTie {
  int step;  // block-local
  int up;    // resolved by the chain: 1 if the tie rounds away from the floor
  long long P;  // inclusive run prefix at the tie (ties rounded down)
};

struct FoldWork {
  double *bsum, *bmin, *bmax, *xstart;  // [3][nblk]
  double *E;                            // [3]
  int *nitems, *nties;                  // [3][nblk]
  FoldItem *items;                      // [3][nblk][kFoldMaxItems]
  FoldTie *ties;                        // [3][nblk][kFoldMaxTies]
  int *status;                          // [3]: 0 ok, 1 overflow/fallback
};

// Device buffers of one waveform slot.
struct DenseBuffers {
  double *knots;     // [17*K]: t | (y,b,c,d) x (e,v,M,S)
  double *e, *v, *M, *S;           // [N+1]
  double *rate[3];                 // gimdot, Phidot, alpdot [N+1]
  double *phase[3];                // gim, Phi, alp [N+1]
  double *h;                       // [2][N+2] time domain hI | hII
  double *t;                       // [N+2]
  int *flags;                      // [4]: [0] non-positive Phidot count
};

// Launch wrappers (aak_kernels.hip).
void launch_rates(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
int fold_exact(const DenseParams &P, const DenseBuffers &B, const FoldWork &W, int nblk_cap, int dir, int nsteps,
               hipStream_t s);
void launch_fold_fill(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
void launch_phidot_guard(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
void launch_modesum(const DenseParams &P, const DenseBuffers &B, hipStream_t s);
void launch_weighted_dot(const double2 *hf, const double2 *data, const double *w, int nbins, double2 *hw_out,
                         double *partials, double *out, hipStream_t s);

}  // namespace aak
