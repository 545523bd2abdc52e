// aak_kernels.hip -- dense stage of the AAK waveform on CDNA4 (gfx950).
//
//   k_rates        spline evaluation of (e, v, M, S) on the dense grid with the
//                  block's knot coefficients staged in LDS, then the three
//                  phase rates from the PN frequency maps (AAK.cc:233-253)
//   fold kernels   the Adams-Bashforth phase folds (AAK.cc:244-293) computed in
//                  parallel but reproducing the sequential FP64 rounding bit for
//                  bit: inside one binade every value is an integer multiple of
//                  the same ulp, so the fold becomes an exact int64 prefix sum;
//                  binade crossings, exact ties and values near zero are
//                  replayed serially in FP64
//   k_modesum      harmonic sum with Bessel amplitudes, Doppler phases, LISA
//                  response and Planck taper (AAK.cc:409-547): per 256-sample
//                  block the amplitudes are Taylor-expanded in e
//                  (k_modesum_setup); blocks where that does not apply are
//                  summed per sample (k_modesum_direct)
//   k_weighted_dot PSD-weighted inner products of a rocFFT spectrum (lengths
//                  the fused FFT of aak_fft.hip does not take), k_dot_final
//                  the last reduction of either form (AAK_manager.cu:352-446)
//
// Compiled with -ffp-contract=off: every expression that feeds the phases
// must round exactly like the reference CPU build.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include <algorithm>
#include <mutex>

#include "aak_dense.hpp"
#include "aak_math.hpp"

namespace aak {

// ------------------------------------------------------------------ rates
constexpr int kRatesThreads = kRatesBlock;
constexpr int kStageCap = 128;  // knot intervals staged in LDS per block (else direct evaluation)

__device__ inline int knot_search(const double *x, int lo, int hi, double t) {
  // largest k in [lo, hi-1] with x[k] <= t (hi is exclusive upper knot)
  while (hi > lo + 1) {
    const int m = (hi + lo) >> 1;
    if (x[m] > t) hi = m;
    else lo = m;
  }
  return lo;
}

__device__ inline double sample_time(const DenseParams &P, int i) {
  if (i == 0) return P.t_first;
  if (i == P.N) return P.t_last;
  return (double)(i - P.i0) * P.dt;
}

// Spline values (e, v, M, S) of one sample straight from the knot table.
__device__ inline void spline_sample(const DenseParams &P, const double *tab, double t, double *y) {
  const KnotGrid &g = P.grid;
  const int K = g.K;
  const double *head = tab + (size_t)kTableArrays * K;
  if (g.flat && t >= knot_time(g, head, K - 1)) {
    // past the table: flat intervals, whose cubic is y[K-1] + 0
#pragma unroll
    for (int a = 0; a < 4; a++) y[a] = tab[(size_t)a * K + K - 1] + 0.0;
    return;
  }
  const int k = knot_index(g, head, t);
  const double x0 = knot_time(g, head, k), x1 = knot_time(g, head, k + 1);
  const double dl = t - x0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double *ya = tab + (size_t)a * K, *ca = tab + (size_t)(4 + a) * K;
    double b, d;
    spline_interval(x0, x1, ya[k], ya[k + 1], ca[k], ca[k + 1], &b, &d);
    y[a] = ya[k] + dl * (b + dl * (ca[k] + dl * d));
  }
}

__global__ __launch_bounds__(kRatesThreads) void k_rates(DenseParams P, DenseBuffers B) {
  __shared__ double s_x[kStageCap + 1];
  __shared__ double s_c[16][kStageCap];
  __shared__ int s_lo, s_cnt;
  const int tid = threadIdx.x;
  // the guard's segment words and ticket start cleared for each waveform
  if (blockIdx.x == 0 && tid <= kGuardGrid) B.guard_work[tid] = (tid < kGuardGrid) ? -2 : 0;
  const int i_lo = blockIdx.x * kRatesThreads;
  if (i_lo > P.N) return;
  // likelihoods read the rates up to the last folded step and e and the
  // amplitude below the end of the signal (<= last_fold + 1); past that a
  // source that plunged needs nothing (single waveforms keep every sample
  // for the phase output and the dense arrays)
  if (!P.full_dense && i_lo > P.last_fold + 1) return;
  const int i_hi = min(i_lo + kRatesThreads - 1, P.N);
  const KnotGrid &g = P.grid;
  const int K = g.K;
  const double *tab = B.knots;
  const double *head = tab + (size_t)kTableArrays * K;
  if (tid == 0) {
    const int klo = knot_index(g, head, sample_time(P, i_lo));
    const int khi = knot_index(g, head, sample_time(P, i_hi));
    s_lo = klo;
    s_cnt = khi - klo + 1;
  }
  __syncthreads();
  const int klo = s_lo, cnt = s_cnt;
  const bool staged = (cnt <= kStageCap);
  if (staged) {
    // one interval per thread: knot time and the cubic of each array
    for (int j = tid; j < cnt; j += kRatesThreads) {
      const int k = klo + j;
      const double x0 = knot_time(g, head, k), x1 = knot_time(g, head, k + 1);
      s_x[j] = x0;
      if (j == cnt - 1) s_x[cnt] = x1;
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const double *ya = tab + (size_t)a * K, *ca = tab + (size_t)(4 + a) * K;
        double b, d;
        spline_interval(x0, x1, ya[k], ya[k + 1], ca[k], ca[k + 1], &b, &d);
        s_c[4 * a][j] = ya[k];
        s_c[4 * a + 1][j] = b;
        s_c[4 * a + 2][j] = ca[k];
        s_c[4 * a + 3][j] = d;
      }
    }
  }
  __syncthreads();
  const int i = i_lo + tid;
  if (i > P.N) return;
  const double t = sample_time(P, i);
  double y[4];
  if (g.flat && t >= knot_time(g, head, K - 1)) {
#pragma unroll
    for (int a = 0; a < 4; a++) y[a] = tab[(size_t)a * K + K - 1] + 0.0;
  } else if (staged) {
    const int k = knot_search(s_x, 0, cnt, t);
    const double dl = t - s_x[k];
#pragma unroll
    for (int a = 0; a < 4; a++)
      y[a] = s_c[4 * a][k] + dl * (s_c[4 * a + 1][k] + dl * (s_c[4 * a + 2][k] + dl * s_c[4 * a + 3][k]));
  } else {
    spline_sample(P, tab, t, y);
  }
  const double e = y[0], v = y[1], M = y[2], S = y[3];
  B.e[i] = e;
  if (P.full_dense) {
    B.v[i] = v;
    B.M[i] = M;
    B.S[i] = S;
  }
  const double Y = P.coslam;
  const double T = PN::dtdm(v, e, Y, S);
  const double R = PN::drdm(v, e, Y, S);
  const double Th = PN::dthetadm(v, e, Y, S);
  const double Ph = PN::dphidm(v, e, Y, S);
  const double phidot = R / T / M;
  B.rate[0][i] = (Th - R) / T / M;
  B.rate[1][i] = phidot;
  B.rate[2][i] = (Ph - Th) / T / M;
  // amplitude of the mode sum (AAK.cc:481), with OmegaPhi of the prograde case
  // from the maps already evaluated here; x^(2/3) as cbrt(x)^2 (within a few
  // ulp of pow: the amplitude scales the waveform and feeds no phase)
  const double Om = (Y > 0) ? Ph / T / M : PN::omega_phi(v, e, Y, S, M);
  const double cr = cbrt(Om * P.Mphys * kMsunSec);
  B.amp[i] = (cr * cr) * P.zeta;
  if (i == 0) B.nu0[0] = R / T / (2. * kPi * M);
  // the reference's test is Phidot <= 0 (AAK.cc:254): a NaN rate is kept as it is
  if (i > P.i0 && i <= P.guard_last && phidot <= 0.) {
    atomicAdd(&B.flags[0], 1);
    atomicMax(&B.flags[5], 0x7fffffff - i);  // the first such sample, for the guard
  }
  // a non-positive rate at t = 0 has no earlier value to take: the waveform is
  // rejected (flags[3]; the likelihood of such a template reads NaN)
  if (i == P.i0 && phidot <= 0.) B.flags[3] = 1;
}

void launch_rates(const DenseParams &P, const DenseBuffers &B, hipStream_t s) {
  const int nb = (P.N + 1 + kRatesThreads - 1) / kRatesThreads;
  hipLaunchKernelGGL(k_rates, dim3(nb), dim3(kRatesThreads), 0, s, P, B);
}

// Non-positive radial rates take the previous (already corrected) value
// (AAK.cc:254): sample i in (i0, guard_last] with Phidot <= 0 gets the rate of
// the last sample j in (i0, i] whose rate is not <= 0 (a NaN rate counts as
// such a sample and is kept, as the reference's test leaves it), or of i0 if
// there is none.  Every sample before the first non-positive one (recorded
// by k_rates) is such a sample, so only [first, guard_last] is walked, split
// into kGuardGrid segments.  Each workgroup takes a ticket, finds its
// segment's last good index and publishes it at once; its carry is the
// maximum over the segments before it, which never wait on anything, so the
// waits are short and cannot form a cycle.  The fix then runs in chunks of
// kGuardThreads x kGuardPer with an inclusive max-scan of the last good index
// over the threads.  Good samples are never written, so a replaced sample
// reads a final value.  The kernel returns at once when k_rates counted no
// such sample, so it is always enqueued and needs no host round trip.
constexpr int kGuardThreads = 1024;

// Look-back words of the single-pass fold (k_fold_resolve), kLW per block and
// phase, reset by k_guard for every waveform: the exact action of a pure
// block (even and odd start) and its grid key, the exact end value, and the
// approximate sum and L1 norm of the block's increments and their inclusive
// prefixes.  "Not yet", "failed" and "not pure" are signalling-NaN patterns
// that arithmetic never produces and no run total or key can equal.
constexpr int kLwA0 = 0, kLwA1 = 1, kLwKey = 2, kLwIncl = 3, kLwAggS = 4, kLwAggL = 5, kLwIncS = 6, kLwIncL = 7;
constexpr long long kLookEmpty = 0x7FF4A5A5A5A5A5A5LL;
constexpr long long kLookFail = 0x7FF4A5A5A5A5A5A6LL;
constexpr long long kLookNone = 0x7FF4A5A5A5A5A5A7LL;
constexpr int kGuardPer = 32;  // consecutive samples per thread and chunk
constexpr int kGuardSpinCap = 1 << 22;

__device__ inline bool phidot_ok(double r) { return !(r <= 0.); }

__global__ __launch_bounds__(kGuardThreads) void k_guard(DenseParams P, DenseBuffers B, FoldReset R) {
  {
    // the fold's look-back words, replay marks and tickets for this waveform
    const int gt = blockIdx.x * blockDim.x + threadIdx.x, gs = gridDim.x * blockDim.x;
    for (int k = gt; k < R.nwords; k += gs) R.words[k] = kLookEmpty;
    if (R.status && gt < 6) R.status[gt] = R.force == 1 ? 1 : 0;
    if (R.ctr && gt < 2) R.ctr[gt] = 0;
  }
  if (P.guard_last <= P.i0 || B.flags[0] == 0) return;
  __shared__ int s_w[kGuardThreads / 64];
  __shared__ int s_t, s_carry;
  double *r = B.rate[1];
  int *gw = B.guard_work;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  if (tid == 0) s_t = atomicAdd(gw + kGuardGrid, 1);
  __syncthreads();
  const int t = s_t;
  const int first = 0x7fffffff - B.flags[5];
  const int n = P.guard_last - first + 1;
  const int per = ((n + kGuardGrid - 1) / kGuardGrid + kGuardThreads - 1) / kGuardThreads * kGuardThreads;
  const int lo = first + t * per, hi = min(P.guard_last, lo + per - 1);
  // the segment's last good index, published at once
  int last = -1;
  for (int i = lo + tid; i <= hi; i += kGuardThreads)
    if (phidot_ok(r[i])) last = i;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) last = max(last, __shfl_xor(last, o, 64));
  if (lane == 0) s_w[wave] = last;
  __syncthreads();
  if (wave == 0) {
    int agg = -1;
    for (int w = 0; w < kGuardThreads / 64; w++) agg = max(agg, s_w[w]);
    if (lane == 0) __hip_atomic_store(gw + t, agg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // carry: the last good index of the segments before this one
    int c = first - 1;
    bool late = false;
    if (lane < t) {
      int v, spins = 0;
      while ((v = __hip_atomic_load(gw + lane, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) == -2) {
        if (++spins > kGuardSpinCap) {
          late = true;
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
      if (!late) c = max(c, v);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) c = max(c, __shfl_xor(c, o, 64));
    const bool any_late = __ballot(late) != 0;
    if (lane == 0) {
      s_carry = c;
      if (any_late) B.flags[3] = 1;  // a wait past the cap rejects the waveform
    }
  }
  __syncthreads();
  int carry = s_carry;
  for (int base = lo; base <= hi; base += kGuardThreads * kGuardPer) {
    const int i0 = base + tid * kGuardPer;
    double v[kGuardPer];
    int lst = -1;
#pragma unroll
    for (int q = 0; q < kGuardPer; q++) {
      const int i = i0 + q;
      v[q] = (i <= hi) ? r[i] : 1.0;
      if (i <= hi && phidot_ok(v[q])) lst = i;
    }
    // inclusive max-scan of the last good index over the threads
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int u = __shfl_up(lst, o, 64);
      if (lane >= o) lst = max(lst, u);
    }
    if (lane == 63) s_w[wave] = lst;
    __syncthreads();
    int before = carry, chunk = carry;
    for (int w = 0; w < kGuardThreads / 64; w++) {
      const int x = s_w[w];
      if (w < wave) before = max(before, x);
      chunk = max(chunk, x);
    }
    // last good index before this thread's samples; its value is read only
    // when the thread's first sample needs it, later ones come from registers
    int src = max(before, __shfl_up(lst, 1, 64));
    if (lane == 0) src = before;
    double srcv = (i0 <= hi && !phidot_ok(v[0])) ? r[src] : 0.0;
#pragma unroll
    for (int q = 0; q < kGuardPer; q++) {
      const int i = i0 + q;
      if (i <= hi) {
        if (phidot_ok(v[q])) srcv = v[q];
        else r[i] = srcv;
      }
    }
    carry = chunk;
    __syncthreads();
  }
}

void launch_phidot_guard(const DenseParams &P, const DenseBuffers &B, const FoldReset &R, hipStream_t s) {
  if (P.guard_last <= P.i0 && R.nwords == 0 && !R.status && !R.ctr) return;
  hipLaunchKernelGGL(k_guard, dim3(kGuardGrid), dim3(kGuardThreads), 0, s, P, B, R);
}

// ------------------------------------------------------------------ folds
__device__ inline double fold_incr(const double *rate, int i0, int dir, int s, double dt) {
  const int i = i0 + dir * s;
  const double r = rate[i];
  const double rm = (s == 0) ? r : rate[i - dir];
  const double d = (1.5 * r - .5 * rm) * dt;
  return dir > 0 ? d : -d;
}

__device__ inline int out_index(const FoldArgs &A, int s) { return A.i0 + A.dir * (s + 1); }

// A block's increments are read with consecutive lanes on consecutive steps
// (coalesced) and staged in LDS; each lane then takes its run of
// kFoldPerThread consecutive steps from LDS.  One pad double per run keeps
// both access patterns free of bank conflicts.
static_assert(kFoldPerThread == 16, "fold_lds_index assumes runs of 16");
constexpr int kFoldLds = kFoldBlock + kFoldBlock / kFoldPerThread;
static_assert(kFoldLds >= 4096, "the serial replay stages 4096 steps in the staging array");
__device__ inline int fold_lds_index(int l) { return l + (l >> 4); }

// Every load is issued before the first use.  An interior block (all its
// steps in range, none of them step 0) reads at fixed offsets from one base
// pointer per lane; other blocks clamp: steps past the end read the last
// step's rates (in bounds) and stage 0.  DIR is the fold direction (A.dir).
template <int DIR>
__device__ inline void stage_increments(const FoldArgs &A, int p, int b, double *sd) {
  const double *rate = A.rate[p];
  double r[kFoldPerThread], rm[kFoldPerThread];
  if (b > 0 && (b + 1) * kFoldBlock <= A.nsteps) {
    const double *rp = rate + (A.i0 + DIR * (b * kFoldBlock + (int)threadIdx.x));
#pragma unroll
    for (int q = 0; q < kFoldPerThread; q++) {
      r[q] = rp[DIR * q * kFoldThreads];
      rm[q] = rp[DIR * q * kFoldThreads - DIR];
    }
  } else {
#pragma unroll
    for (int q = 0; q < kFoldPerThread; q++) {
      const int s = min(b * kFoldBlock + (int)threadIdx.x + q * kFoldThreads, A.nsteps - 1);
      const int i = A.i0 + DIR * s;
      r[q] = rate[i];
      rm[q] = rate[s == 0 ? i : i - DIR];
    }
  }
#pragma unroll
  for (int q = 0; q < kFoldPerThread; q++) {
    const int l = threadIdx.x + q * kFoldThreads;
    const double d = (1.5 * r[q] - .5 * rm[q]) * A.dt;
    sd[fold_lds_index(l)] = (b * kFoldBlock + l < A.nsteps) ? (DIR > 0 ? d : -d) : 0.0;
  }
  __syncthreads();
}

// Block-wide exclusive scans over kFoldThreads lanes: inclusive shuffle scan
// inside each 64-lane wave, wave totals through LDS.  `sh` needs
// kFoldThreads / 64 entries.  Integer scans are exact; the double scan is an
// approximation consumer (the fold's binade prediction) whose error the
// bound E covers.
template <class T>
__device__ inline T wave_incl_sum(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

template <class T>
__device__ inline T block_excl_sum(T v, T *sh, T *total) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const T incl = wave_incl_sum(v);
  if (lane == 63) sh[wave] = incl;
  __syncthreads();
  T before = T(0), tot = T(0);
#pragma unroll
  for (int w = 0; w < kFoldThreads / 64; w++) {
    const T x = sh[w];
    if (w < wave) before += x;
    tot += x;
  }
  if (total) *total = tot;
  __syncthreads();
  return before + incl - v;
}

// Segmented inclusive scan of (flag, value) pairs over the block, lanes in
// order: a set flag starts a new segment.  Returns the inclusive value of the
// lane's segment prefix; shv/shf need kFoldThreads / 64 entries.
__device__ inline long long block_seg_incl(long long v, int f, long long *shv, int *shf) {
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const long long uv = __shfl_up(v, o, 64);
    const int uf = __shfl_up(f, o, 64);
    if (lane >= o && !f) {
      v += uv;
      f |= uf;
    }
  }
  if (lane == 63) {
    shv[wave] = v;
    shf[wave] = f;
  }
  __syncthreads();
  if (!f) {
    // carry in from the earlier waves, back to the nearest segment start
    for (int w = wave - 1; w >= 0; w--) {
      v += shv[w];
      if (shf[w]) break;
    }
  }
  __syncthreads();
  return v;
}

// approximate block sums and prefix ranges
template <int DIR>
__global__ __launch_bounds__(kFoldThreads) void k_fold_sums(FoldArgs A, FoldWork W) {
  __shared__ double sh[2 * (kFoldThreads / 64)];
  __shared__ double sd[kFoldLds];
  const int p = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  const int base = b * kFoldBlock + tid * kFoldPerThread;
  stage_increments<DIR>(A, p, b, sd);
  double run = 0.0, mn = 0.0, mx = 0.0;
  for (int j = 0; j < kFoldPerThread; j++) {
    const int s = base + j;
    if (s < A.nsteps) {
      run += sd[fold_lds_index(tid * kFoldPerThread + j)];
      mn = fmin(mn, run);
      mx = fmax(mx, run);
    }
  }
  double tot;
  const double off = block_excl_sum(run, sh, &tot);
  double lmn = off + mn, lmx = off + mx;
  for (int o = 32; o > 0; o >>= 1) {
    lmn = fmin(lmn, __shfl_xor(lmn, o, 64));
    lmx = fmax(lmx, __shfl_xor(lmx, o, 64));
  }
  if ((tid & 63) == 0) {
    sh[tid >> 6] = lmn;
    sh[kFoldThreads / 64 + (tid >> 6)] = lmx;
  }
  __syncthreads();
  if (tid == 0) {
    double bmn = sh[0], bmx = sh[kFoldThreads / 64];
    for (int w = 1; w < kFoldThreads / 64; w++) {
      bmn = fmin(bmn, sh[w]);
      bmx = fmax(bmx, sh[kFoldThreads / 64 + w]);
    }
    W.bsum[p * A.nblk + b] = tot;
    W.bmin[p * A.nblk + b] = bmn;
    W.bmax[p * A.nblk + b] = bmx;
  }
}

// approximate block start values and the error bound E
__global__ __launch_bounds__(1024) void k_fold_plan(FoldArgs A, FoldWork W) {
  __shared__ double s_w[16], s_m[16];
  const int p = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int per = (A.nblk + 1023) / 1024;
  double loc = 0.0;
  for (int j = 0; j < per; j++) {
    const int b = tid * per + j;
    if (b < A.nblk) loc += W.bsum[p * A.nblk + b];
  }
  // exclusive prefix of loc over the threads: shuffles inside each wave, then
  // the wave totals (approximate starts: the bound E covers their rounding)
  double incl = loc;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(incl, o, 64);
    if (lane >= o) incl += u;
  }
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  double before = 0.0;
  for (int w = 0; w < wave; w++) before += s_w[w];
  const double up = __shfl_up(incl, 1, 64);
  double run = A.x0[p] + (before + (lane > 0 ? up : 0.0));
  double amax = fabs(A.x0[p]);
  for (int j = 0; j < per; j++) {
    const int b = tid * per + j;
    if (b < A.nblk) {
      W.xstart[p * A.nblk + b] = run;
      amax = fmax(amax, fmax(fabs(run + W.bmin[p * A.nblk + b]), fabs(run + W.bmax[p * A.nblk + b])));
      run += W.bsum[p * A.nblk + b];
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) amax = fmax(amax, __shfl_xor(amax, o, 64));
  if (lane == 0) s_m[wave] = amax;
  __syncthreads();
  if (tid == 0) {
    double m = s_m[0];
    for (int w = 1; w < 16; w++) m = fmax(m, s_m[w]);
    int ex = 0;
    frexp(m > 0.0 ? m : 1.0, &ex);
    W.E[p] = ((double)A.nsteps + 2.0 * kFoldBlock) * ldexp(1.0, ex - 52);
    W.status[p] = W.force_replay == 1 ? 1 : 0;
  }
}

__device__ inline int binade_key(double a, double E) {
  const double aa = fabs(a);
  const double lo = aa - E, hi = aa + E;
  if (!(lo > 0.0)) return -1;
  int klo, khi;
  frexp(lo, &klo);
  frexp(hi, &khi);
  if (klo != khi) return -1;
  return (klo + 4096) * 2 + (a > 0.0 ? 1 : 0);
}

// Fallback: plain serial replay of one phase's whole fold, by one workgroup
// with buf (bufn doubles of LDS) as staging; used when the item bookkeeping
// overflowed or a check failed (correct for any input).
__device__ void fold_serial_replay(const FoldArgs &A, const FoldWork &W, int p, double *buf, int bufn) {
  if (threadIdx.x == 0) atomicAdd(W.fallbacks, 1);
  double x = A.x0[p];
  if (threadIdx.x == 0) A.out[p][A.i0] = x;
  for (int base = 0; base < A.nsteps; base += bufn) {
    const int n = min(bufn, A.nsteps - base);
    for (int j = threadIdx.x; j < n; j += blockDim.x) buf[j] = fold_incr(A.rate[p], A.i0, A.dir, base + j, A.dt);
    __syncthreads();
    if (threadIdx.x == 0) {
      for (int j = 0; j < n; j++) {
        x = x + buf[j];
        buf[j] = x;
      }
    }
    __syncthreads();
    for (int j = threadIdx.x; j < n; j += blockDim.x) A.out[p][out_index(A, base + j)] = buf[j];
    __syncthreads();
  }
}

// Classify steps, build the block's item list (PLAN) or write the exact
// fold values of run steps (WRITE, after the chain resolved the run starts).
template <bool WRITE, int DIR>
__global__ __launch_bounds__(kFoldThreads, WRITE ? 3 : 4) void k_fold_local(FoldArgs A, FoldWork W) {
  __shared__ double shd[kFoldThreads / 64];
  __shared__ long long shl[2 * (kFoldThreads / 64)];
  __shared__ int shi[kFoldThreads];
  __shared__ int shf[kFoldThreads / 64];
  __shared__ double sd[kFoldLds];
  const int p = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  if (W.status[p]) {
    // the write pass of a phase the chain could not resolve: its first
    // workgroup replays the fold serially
    if (WRITE && b == 0) fold_serial_replay(A, W, p, sd, 4096);
    return;
  }
  const int base = b * kFoldBlock + tid * kFoldPerThread;
  const double E = W.E[p];
  stage_increments<DIR>(A, p, b, sd);
  double d[kFoldPerThread];
  double tsum = 0.0;
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    d[j] = sd[fold_lds_index(tid * kFoldPerThread + j)];
    tsum += d[j];
  }
  double dummy;
  double a0 = W.xstart[p * A.nblk + b] + block_excl_sum(tsum, shd, &dummy);
  // keys: >= 0 binade grid, -1 exact replay, -2 beyond the end
  int key[kFoldPerThread];
  long long r[kFoldPerThread];
  int tie_mask = 0, head_mask = 0;
  int prev_key = -3;
  // key of the previous lane's last step, for head detection across lanes
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    const int s = base + j;
    if (s >= A.nsteps) {
      key[j] = -2;
      r[j] = 0;
      continue;
    }
    const double a1 = a0 + d[j];
    const int k0 = binade_key(a0, E);
    const int k1 = binade_key(a1, E);
    if (k0 >= 0 && k0 == k1) {
      key[j] = k0;
      const int kexp = (k0 >> 1) - 4096;
      const double f = ldexp(d[j], 53 - kexp);
      const double fl = floor(f);
      const double fr = f - fl;
      long long ri = (long long)fl;
      if (fr > 0.5) ri += 1;
      else if (fr == 0.5) tie_mask |= (1 << j);
      r[j] = ri;
    } else {
      key[j] = -1;
      r[j] = 0;
    }
    a0 = a1;
  }
  // head flags need the previous lane's last key
  shi[tid] = key[kFoldPerThread - 1];
  __syncthreads();
  prev_key = (tid > 0) ? shi[tid - 1] : -3;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    const int kprev = (j == 0) ? prev_key : key[j - 1];
    const bool h = (key[j] != -2) && (key[j] == -1 || kprev == -1 || kprev == -3 || key[j] != kprev);
    if (h) head_mask |= (1 << j);
  }
  // does the next lane start a new item (or lie beyond the end)?
  shi[tid] = ((head_mask & 1) || key[0] == -2) ? 1 : 0;
  __syncthreads();
  const bool next_starts = (tid + 1 < kFoldThreads) ? (shi[tid + 1] != 0) : true;
  __syncthreads();
  // segmented carry of r across lanes: summary (has_head, tail_sum)
  long long tail = 0;
  int has_head = 0;
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    if (head_mask & (1 << j)) {
      has_head = 1;
      tail = 0;
    }
    tail += r[j];
  }
  // carry into the lane: the segmented inclusive scan of the previous lane
  const long long seg_incl = block_seg_incl(tail, has_head, shl, shf);
  long long carry = __shfl_up(seg_incl, 1, 64);
  if ((tid & 63) == 63) shl[kFoldThreads / 64 + (tid >> 6)] = seg_incl;  // for the next wave's lane 0
  __syncthreads();
  if ((tid & 63) == 0) carry = (tid > 0) ? shl[kFoldThreads / 64 + (tid >> 6) - 1] : 0;
  __syncthreads();
  // item and tie numbering
  const int nh = __popc(head_mask);
  int nheads_tot;
  const int item_base = block_excl_sum(nh, shi, &nheads_tot);
  const int nt = __popc(tie_mask);
  int nties_tot;
  const int tie_base = block_excl_sum(nt, shi, &nties_tot);
  if (!WRITE && tid == 0) {
    W.nitems[p * A.nblk + b] = nheads_tot;
    W.nties[p * A.nblk + b] = nties_tot;
    if (nheads_tot > kFoldMaxItems || nties_tot > kFoldMaxTies) atomicExch(&W.status[p], 1);
  }
  if (nheads_tot > kFoldMaxItems || nties_tot > kFoldMaxTies) return;
  FoldItem *items = W.items + ((size_t)p * A.nblk + b) * kFoldMaxItems;
  FoldTie *ties = W.ties + ((size_t)p * A.nblk + b) * kFoldMaxTies;
  long long P = carry;
  int item = item_base - 1, tcount = tie_base;
  int c_item = -1, c_k = 0, c_tf = 0;  // WRITE: the item of the current run
  long long c_m = 0;
  double outv[kFoldPerThread];  // WRITE: values of the run steps, NaN elsewhere
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) outv[j] = __builtin_nan("");
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    if (key[j] == -2) break;
    const int sl = tid * kFoldPerThread + j;
    if (head_mask & (1 << j)) {
      item++;
      P = 0;
      if (!WRITE) {
        items[item].kind = (key[j] == -1) ? 1 : 0;
        items[item].k = (key[j] >= 0) ? (key[j] >> 1) - 4096 : 0;
        items[item].sgn = (key[j] >= 0) ? ((key[j] & 1) ? 1 : -1) : 0;
        items[item].first = sl;
        items[item].tie_first = tcount;
        items[item].d = d[j];
      }
    }
    P += r[j];
    const bool last_of_run =
        (key[j] >= 0) && ((j + 1 < kFoldPerThread) ? (((head_mask >> (j + 1)) & 1) || key[j + 1] == -2) : next_starts);
    if (tie_mask & (1 << j)) {
      if (!WRITE) {
        ties[tcount].step = sl;
        ties[tcount].P = P;
        ties[tcount].up = 0;
      }
      tcount++;
    }
    if (!WRITE) {
      // the run total is the inclusive prefix at the run's last step
      if (last_of_run) items[item].R = P;
    } else if (key[j] >= 0) {
      if (item != c_item) {
        c_item = item;
        c_m = items[item].m_start;
        c_k = items[item].k;
        c_tf = items[item].tie_first;
      }
      long long ups = 0;
      for (int q = c_tf; q < tcount; q++) ups += ties[q].up;
      const long long m = c_m + P + ups;
      outv[j] = ldexp((double)m, c_k - 53);
    }
  }
  if (WRITE) {
    // coalesced stores of the run steps; single steps were written by the chain
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) sd[fold_lds_index(tid * kFoldPerThread + j)] = outv[j];
    __syncthreads();
    // step s goes to A.i0 + DIR (s + 1)
    double *op = A.out[p] + (A.i0 + DIR * (b * kFoldBlock + tid + 1));
#pragma unroll
    for (int q = 0; q < kFoldPerThread; q++) {
      const int s = b * kFoldBlock + tid + q * kFoldThreads;
      const double v = sd[fold_lds_index(tid + q * kFoldThreads)];
      if (s < A.nsteps && !__builtin_isnan(v)) op[DIR * q * kFoldThreads] = v;
    }
  }
}

// Exact start of every run, resolution of the exact ties, FP64 replay of the
// single steps near binade boundaries.  The item and tie lists of all blocks
// are gathered into LDS; consecutive runs on one binade grid without ties
// form a segment whose total is a segmented int64 scan, so the one serial
// lane only visits segment heads (binade changes, ties, single steps), each
// packed beforehand into one walk record; the run starts inside a segment
// are then filled in parallel, one item per lane, and re-checked against
// their binade.  Lists too long for LDS take a serial walk over global
// memory with the same arithmetic.
constexpr int kChainThreads = 256;
constexpr int kChainMaxBlocks = 2048;
constexpr int kChainMaxItems = 1152;
constexpr int kChainMaxTies = 512;

// One item (run or single step) in walk order; returns false on a failed
// binade check.  ties: the item's ties, nt of them.
template <class ItemRef, class TiePtr>
__device__ inline bool chain_item(const FoldArgs &A, int p, int b, ItemRef I, TiePtr ties, int nt, double &x,
                                  long long extra_R) {
  if (I.kind == 1) {
    x = x + I.d;
    A.out[p][out_index(A, b * kFoldBlock + I.first)] = x;
    return true;
  }
  int kx;
  frexp(x, &kx);
  if (kx != I.k || (x > 0.0 ? 1 : -1) != I.sgn) return false;
  const long long m = (long long)ldexp(x, 53 - I.k);
  I.m_start = m;
  long long ups = 0;
  for (int q = 0; q < nt; q++) {
    const long long val = m + ties[q].P + ups;
    const int up = (int)(val & 1LL);
    ties[q].up = up;
    ups += up;
  }
  const long long mend = m + I.R + extra_R + ups;
  const long long am = mend < 0 ? -mend : mend;
  if (am < (1LL << 52) || am >= (1LL << 53)) return false;
  x = ldexp((double)mend, I.k - 53);
  return true;
}

static_assert(kChainThreads == kFoldThreads, "the chain uses the fold's block scans");

// One serial step of the chain's walk: a segment head (run or single step).
struct ChainHead {
  union {
    long long R;  // kind 0: total of the segment's runs (ties rounded down)
    double d;     // kind 1: the step's increment
  };
  int out;        // kind 1: output index of the step
  short k, tie_first, nt;
  signed char kind, sgn;
};

// Exclusive prefix of the per-thread chunk sums of v[0..n) (chunks of
// ceil(n / kChainThreads) consecutive entries), written in place; returns the
// total.  One wave-level scan and one block barrier pair.
__device__ inline int chain_excl_scan(int *v, int n, int *sh) {
  const int tid = threadIdx.x;
  const int per = (n + kChainThreads - 1) / kChainThreads;
  const int lo = min(n, tid * per), hi = min(n, lo + per);
  int sum = 0;
  for (int j = lo; j < hi; j++) sum += v[j];
  int total;
  int run = block_excl_sum(sum, sh, &total);
  for (int j = lo; j < hi; j++) {
    const int c = v[j];
    v[j] = run;
    run += c;
  }
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kChainThreads) void k_fold_chain(FoldArgs A, FoldWork W) {
  __shared__ int s_io[kChainMaxBlocks + 1], s_to[kChainMaxBlocks + 1];
  __shared__ int s_tmp[kChainThreads / 64];
  __shared__ long long s_ltmp[2 * (kChainThreads / 64)];
  __shared__ int s_ftmp[kChainThreads / 64];
  __shared__ FoldItem s_items[kChainMaxItems];
  __shared__ FoldTie s_ties[kChainMaxTies];
  __shared__ long long s_pref[kChainMaxItems];  // exclusive R prefix inside the segment
  __shared__ int s_head[kChainMaxItems];        // head flags
  __shared__ int s_seg[kChainMaxItems];         // item -> block (gather), then item -> segment
  __shared__ int s_hlist[kChainMaxItems];       // segment -> head item
  __shared__ long long s_mh[kChainMaxItems];    // segment -> exact start of the head run
  __shared__ ChainHead s_hinfo[kChainMaxItems];
  __shared__ int s_nheads, s_ok;
  const int p = blockIdx.x;
  const int tid = threadIdx.x;
  if (W.status[p]) return;
  const int nblk = A.nblk;
  const int *gni = W.nitems + p * nblk, *gnt = W.nties + p * nblk;
  FoldItem *gitems = W.items + (size_t)p * nblk * kFoldMaxItems;
  FoldTie *gties = W.ties + (size_t)p * nblk * kFoldMaxTies;
  bool fits = nblk <= kChainMaxBlocks;
  int TI = 0, TT = 0;
  if (fits) {
    for (int b = tid; b < nblk; b += kChainThreads) {
      s_io[b] = gni[b];
      s_to[b] = gnt[b];
    }
    __syncthreads();
    TI = chain_excl_scan(s_io, nblk, s_tmp);
    TT = chain_excl_scan(s_to, nblk, s_tmp);
    fits = TI <= kChainMaxItems && TT <= kChainMaxTies;
  }
  if (!fits) {
    if (tid != 0) return;
    double x = A.x0[p];
    A.out[p][A.i0] = x;
    for (int b = 0; b < nblk; b++) {
      FoldItem *it = gitems + (size_t)b * kFoldMaxItems;
      FoldTie *tb = gties + (size_t)b * kFoldMaxTies;
      const int ni = gni[b], nt = gnt[b];
      for (int q = 0; q < ni; q++) {
        const int t0 = it[q].tie_first, t1 = (q + 1 < ni) ? it[q + 1].tie_first : nt;
        if (!chain_item<FoldItem &>(A, p, b, it[q], tb + t0, t1 - t0, x, 0)) {
          W.status[p] = 1;
          return;
        }
      }
    }
    return;
  }
  if (tid == 0) {
    s_io[nblk] = TI;
    s_to[nblk] = TT;
  }
  __syncthreads();
  // block of every item (LDS only), and the ties of every block
  for (int b = tid; b < nblk; b += kChainThreads) {
    const int io = s_io[b], ni = s_io[b + 1] - io;
    for (int j = 0; j < ni; j++) s_seg[io + j] = b;
    const int to = s_to[b], nt = s_to[b + 1] - to;
    for (int j = 0; j < nt; j++) s_ties[to + j] = gties[(size_t)b * kFoldMaxTies + j];
  }
  __syncthreads();
  // gather: independent loads, items get their block and a block-global tie offset
  for (int i = tid; i < TI; i += kChainThreads) {
    const int b = s_seg[i];
    FoldItem it = gitems[(size_t)b * kFoldMaxItems + (i - s_io[b])];
    it.blk = b;
    it.tie_first += s_to[b];
    s_items[i] = it;
  }
  __syncthreads();
  // tie count of item i: up to the next item's first tie, or the block's end
  auto nties_of = [&](int i) {
    const FoldItem &I = s_items[i];
    const int end = (i + 1 < TI && s_items[i + 1].blk == I.blk) ? s_items[i + 1].tie_first : s_to[I.blk + 1];
    return end - I.tie_first;
  };
  // segment heads: single steps, items with ties and their successors,
  // changes of binade or sign
  for (int i = tid; i < TI; i += kChainThreads) {
    const FoldItem &I = s_items[i];
    int h = 1;
    if (i > 0) {
      const FoldItem &Q = s_items[i - 1];
      h = (I.kind != 0 || Q.kind != 0 || nties_of(i) > 0 || nties_of(i - 1) > 0 || I.k != Q.k || I.sgn != Q.sgn) ? 1 : 0;
    }
    s_head[i] = h;
  }
  __syncthreads();
  const int per = (TI + kChainThreads - 1) / kChainThreads;
  const int lo = min(TI, tid * per), hi = min(TI, lo + per);
  {
    // segmented exclusive scan of R over items: chunk summaries, then lanes
    long long tail = 0;
    int has = 0, cnt = 0;
    for (int j = lo; j < hi; j++) {
      if (s_head[j]) {
        has = 1;
        tail = 0;
        cnt++;
      }
      tail += s_items[j].kind == 0 ? s_items[j].R : 0;
    }
    const long long incl = block_seg_incl(tail, has, s_ltmp, s_ftmp);
    long long carry = __shfl_up(incl, 1, 64);
    if ((tid & 63) == 63) s_ltmp[kChainThreads / 64 + (tid >> 6)] = incl;
    __syncthreads();
    if ((tid & 63) == 0) carry = (tid > 0) ? s_ltmp[kChainThreads / 64 + (tid >> 6) - 1] : 0;
    int nheads;
    int pos = block_excl_sum(cnt, s_tmp, &nheads);
    long long run = carry;
    for (int j = lo; j < hi; j++) {
      if (s_head[j]) {
        run = 0;
        s_hlist[pos++] = j;
      }
      s_pref[j] = run;
      s_seg[j] = pos - 1;
      run += s_items[j].kind == 0 ? s_items[j].R : 0;
    }
    if (tid == 0) s_nheads = nheads;
  }
  __syncthreads();
  const int nh = s_nheads;
  // one walk record per segment
  for (int q = tid; q < nh; q += kChainThreads) {
    const int h = s_hlist[q];
    const int last = ((q + 1 < nh) ? s_hlist[q + 1] : TI) - 1;
    const FoldItem &I = s_items[h];
    ChainHead H;
    H.kind = (signed char)I.kind;
    H.sgn = (signed char)I.sgn;
    H.k = (short)I.k;
    H.tie_first = (short)I.tie_first;
    H.nt = (short)nties_of(h);
    if (I.kind == 1) {
      H.d = I.d;
      H.out = out_index(A, I.blk * kFoldBlock + I.first);
    } else {
      H.R = (last > h) ? s_pref[last] + s_items[last].R : I.R;
      H.out = 0;
    }
    s_hinfo[q] = H;
  }
  __syncthreads();
  // serial walk over the segment heads; the next record is read ahead
  if (tid == 0) {
    double x = A.x0[p];
    A.out[p][A.i0] = x;
    int ok = 1;
    ChainHead Hn = s_hinfo[0];
    for (int q = 0; q < nh; q++) {
      const ChainHead H = Hn;
      if (q + 1 < nh) Hn = s_hinfo[q + 1];
      if (H.kind == 1) {
        x = x + H.d;
        A.out[p][H.out] = x;
        s_mh[q] = 0;
        continue;
      }
      int kx;
      frexp(x, &kx);
      if (kx != H.k || (x > 0.0 ? 1 : -1) != H.sgn) {
        ok = 0;
        break;
      }
      const long long m = (long long)ldexp(x, 53 - H.k);
      s_mh[q] = m;
      long long ups = 0;
      for (int t = 0; t < H.nt; t++) {
        FoldTie &T = s_ties[H.tie_first + t];
        const int up = (int)((m + T.P + ups) & 1LL);
        T.up = up;
        ups += up;
      }
      const long long mend = m + H.R + ups;
      const long long am = mend < 0 ? -mend : mend;
      if (am < (1LL << 52) || am >= (1LL << 53)) {
        ok = 0;
        break;
      }
      x = ldexp((double)mend, H.k - 53);
    }
    s_ok = ok;
    if (!ok) W.status[p] = 1;
  }
  __syncthreads();
  if (!s_ok) return;
  // exact start of every run, re-checked against its binade
  int bad = 0;
  for (int i = tid; i < TI; i += kChainThreads) {
    const int q = s_seg[i];
    long long m = s_mh[q];
    if (i > s_hlist[q]) {
      m += s_pref[i];
      const long long am = m < 0 ? -m : m;
      if (am < (1LL << 52) || am >= (1LL << 53)) bad = 1;
    }
    s_items[i].m_start = m;
  }
  if (bad) W.status[p] = 1;
  __syncthreads();
  for (int i = tid; i < TI; i += kChainThreads) {
    const int b = s_items[i].blk;
    gitems[(size_t)b * kFoldMaxItems + (i - s_io[b])].m_start = s_items[i].m_start;
  }
  for (int b = tid; b < nblk; b += kChainThreads) {
    const int to = s_to[b], nt = s_to[b + 1] - to;
    for (int j = 0; j < nt; j++) gties[(size_t)b * kFoldMaxTies + j].up = s_ties[to + j].up;
  }
}

// Single-pass alternative to k_fold_local<false> + k_fold_chain +
// k_fold_local<true>: each workgroup classifies its block as above, keeps the
// item and tie lists in LDS, and obtains the exact value before its first
// step by a decoupled look-back over the blocks before it (Merrill and
// Garland, "Single-pass Parallel Prefix Scan with Decoupled Look-back",
// 2016), with the exact fold as the operator:
// - a block that is one run on one binade grid without ties ("pure") acts on
//   the value before it as m -> m + R (m the value in units of its binade's
//   ulp), so consecutive pure blocks on one grid compose by adding their
//   totals; it publishes R as soon as it has classified its steps.  Its grid
//   is that of its first step, which every block can compute from the plan's
//   approximate block start, so R is published alone;
// - every block publishes its exact end value after its own serial walk.
// Each published value is one 64-bit word written and polled with relaxed
// device-scope atomics, so no fence orders it against anything else; "not
// yet" and "failed" are signalling-NaN patterns arithmetic never produces.
// The look-back reads 64 predecessors at a time with one wave, adds the
// totals of the nearest run of pure blocks on one grid and stops at the first
// block with an exact end value; it waits only on blocks that are not pure on
// that grid.  Every block checks its own start and end against its binades
// as the chain does, so a result that passes is the chain's result bit for
// bit; a failed check, an overflowing list or a wait longer than kLookSpinCap
// polls marks the phase, which k_fold_replay then replays serially.
// Blocks take tickets from a counter in the order they start, so every block
// a workgroup waits on has started before it and waits only on earlier
// blocks itself: the waits cannot form a cycle.
constexpr int kLookSpinCap = 1 << 18;

__device__ inline long long look_load(const long long *v) {
  return __hip_atomic_load(v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void look_store(long long *v, long long x) {
  __hip_atomic_store(v, x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// a value every lane holds, kept in scalar registers
__device__ inline double uniform_double(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// a 64-bit value of lane q, read by every lane of the wave
__device__ inline long long lane_read64(long long v, int q) {
  const int lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffLL), q);
  const int hi = __builtin_amdgcn_readlane((int)(v >> 32), q);
  return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// binade key of an exact value, as binade_key with E = 0 (-1 for 0, inf, NaN)
__device__ inline int exact_key(double x) {
  if (!(fabs(x) > 0.0) || !(fabs(x) < __builtin_inf())) return -1;
  int kx;
  frexp(x, &kx);
  return (kx + 4096) * 2 + (x > 0.0 ? 1 : 0);
}

// Look-back state of block j: 2 end value known (or failed), 1 pure with its
// action published, 0 neither yet (or not pure).  The action of a pure block
// with ties on its grid depends on the parity of the value before it (each
// tie rounds to even), so it is published as two totals: a[q] is added to m
// when m is even (q = 0) or odd (q = 1).  Such actions compose: m -> m +
// F[m & 1] followed by m -> m + G[m & 1] is m -> m + H[m & 1] with
// H[q] = F[q] + G[(q + F[q]) & 1].
struct LookState {
  int st;
  long long v;       // st 2: end value bits
  long long a0, a1;  // st 1: the totals for an even and an odd start
};
__device__ inline LookState look_state(const long long *lw, int j, int &key) {
  const long long *w = lw + (size_t)j * kLW;
  const long long iw = look_load(w + kLwIncl);  // all loads in flight at once
  const long long w0 = look_load(w + kLwA0), w1 = look_load(w + kLwA1), kw = look_load(w + kLwKey);
  key = -1;
  if (iw != kLookEmpty) return {2, iw, 0, 0};
  if (w0 != kLookEmpty && w0 != kLookNone && w1 != kLookEmpty && kw != kLookEmpty) {
    key = (int)kw;
    return {1, 0, w0, w1};
  }
  return {0, 0, 0, 0};
}

// (far then near) as one action
__device__ inline void act_compose(long long f0, long long f1, long long &n0, long long &n1) {
  const long long h0 = f0 + ((f0 & 1) ? n1 : n0);
  const long long h1 = f1 + ((f1 & 1) ? n0 : n1);
  n0 = h0;
  n1 = h1;
}

// The value before a block from an exact end value w and the pure blocks
// combined after it (key K, action c0/c1; K < 0: none).
__device__ inline bool look_finish(long long w, int K, long long c0, long long c1, double &x_out) {
  if (w == kLookFail) return false;
  double x = __longlong_as_double(w);
  if (K >= 0) {
    if (exact_key(x) != K) return false;
    const int k = (K >> 1) - 4096;
    long long m = (long long)ldexp(x, 53 - k);
    m += (m & 1) ? c1 : c0;
    const long long am = m < 0 ? -m : m;
    if (am < (1LL << 52) || am >= (1LL << 53)) return false;
    x = ldexp((double)m, k - 53);
  }
  x_out = x;
  return true;
}

// Wave 0: exact value before block b's first step.  Returns false when the
// phase failed (a failed predecessor, a failed binade check, a wait past the
// cap).
__device__ inline bool fold_look_back(const FoldArgs &A, const long long *lw, int p, int b, double &x_out) {
  const int lane = threadIdx.x & 63;
  int j = b - 1;              // nearest block not yet combined
  int K = -1;                 // binade key of the combined pure blocks
  long long c0 = 0, c1 = 0;   // their combined action
  int spins = 0;
  for (;;) {
    const int jb = j - lane;
    // block -1 stands for the start value x0
    LookState S = {2, __double_as_longlong(A.x0[p]), 0, 0};
    int key = -1;
    if (jb >= 0) S = look_state(lw, jb, key);
    const int s0 = __shfl(S.st, 0, 64);
    const int k0 = __shfl(key, 0, 64);
    const int Kc = (K >= 0) ? K : (s0 == 1 ? k0 : -2);
    const bool ext = S.st == 1 && key == Kc && key >= 0;
    const unsigned long long mask = __ballot(ext);
    const int s = (~mask) ? __builtin_ctzll(~mask) : 64;
    // ordered composition of lanes 0..s-1 (lane 0 nearest): after the step
    // with offset o, lane l (a multiple of 2o) holds lanes l..l+2o-1
    long long n0 = (lane < s) ? S.a0 : 0, n1 = (lane < s) ? S.a1 : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const long long f0 = __shfl_down(n0, o, 64), f1 = __shfl_down(n1, o, 64);
      if (lane + o < 64) act_compose(f0, f1, n0, n1);
    }
    {
      // the window lies before the blocks combined so far
      long long w0 = __shfl(n0, 0, 64), w1 = __shfl(n1, 0, 64);
      act_compose(w0, w1, c0, c1);
    }
    if (s > 0) K = Kc;
    if (s == 64) {
      j -= 64;
      continue;
    }
    const int js = j - s;
    const int ss = __shfl(S.st, s, 64);
    if (ss == 2) return look_finish(__shfl(S.v, s, 64), K, c0, c1, x_out);
    // not yet published, or pure on another grid: wait for a change of state;
    // an end value is used at once, a newly published action is combined by
    // the next window, which starts at that block
    LookState T;
    int tk;
    do {
      if (++spins > kLookSpinCap) return false;
      __builtin_amdgcn_s_sleep(1);
      T = look_state(lw, js, tk);
    } while (T.st == ss);
    if (T.st == 2) return look_finish(T.v, K, c0, c1, x_out);
    j = js;
  }
}

// Wave 0: the approximate value before block b (S) and |x0| plus the L1 norm
// of the increments before it (L), from the sums every block publishes right
// after loading its rates; the nearest block with published prefixes ends
// the walk back.  The blocks waited on publish their sums without waiting
// themselves, so the waits are short.  Returns false after a wait past the
// cap.
__device__ inline bool approx_look_back(const FoldArgs &A, const long long *lw, int p, int b, double &S, double &L) {
  const int lane = threadIdx.x & 63;
  int j = b - 1;
  double s = 0.0, l = 0.0;
  int spins = 0;
  for (;;) {
    const int jb = j - lane;
    int st = 2;  // block -1 stands for the start value
    double vs = A.x0[p], vl = fabs(A.x0[p]);
    if (jb >= 0) {
      const long long *w = lw + (size_t)jb * kLW;
      const long long is = look_load(w + kLwIncS), il = look_load(w + kLwIncL);
      const long long as = look_load(w + kLwAggS), al = look_load(w + kLwAggL);
      if (is != kLookEmpty && il != kLookEmpty) {
        vs = __longlong_as_double(is);
        vl = __longlong_as_double(il);
      } else if (as != kLookEmpty && al != kLookEmpty) {
        st = 1;
        vs = __longlong_as_double(as);
        vl = __longlong_as_double(al);
      } else {
        st = 0;
      }
    }
    const unsigned long long m2 = __ballot(st == 2), m0 = __ballot(st == 0);
    const int f2 = m2 ? __builtin_ctzll(m2) : 64, f0 = m0 ? __builtin_ctzll(m0) : 64;
    // lanes before the first unpublished block, or through the nearest prefix
    const int upto = (f0 < f2) ? f0 : (f2 < 64 ? f2 + 1 : 64);
    double ts = (lane < upto) ? vs : 0.0, tl = (lane < upto) ? vl : 0.0;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ts += __shfl_xor(ts, o, 64);
      tl += __shfl_xor(tl, o, 64);
    }
    s += ts;
    l += tl;
    if (f0 < f2) {
      j -= f0;
      for (;;) {
        if (++spins > kLookSpinCap) return false;
        __builtin_amdgcn_s_sleep(1);
        const long long *w = lw + (size_t)j * kLW;
        const long long as = look_load(w + kLwAggS), al = look_load(w + kLwAggL);
        if (as != kLookEmpty && al != kLookEmpty) break;
      }
      continue;
    }
    if (f2 < 64) {
      S = s;
      L = l;
      return true;
    }
    j -= 64;
  }
}

template <int DIR>
__global__ __launch_bounds__(kFoldThreads, 3) void k_fold_resolve(FoldArgs A, FoldWork W) {
  __shared__ double shd[kFoldThreads / 64];
  __shared__ long long shl[2 * (kFoldThreads / 64)];
  __shared__ int shi[kFoldThreads];
  __shared__ int shf[kFoldThreads / 64];
  __shared__ double sd[kFoldLds];
  __shared__ FoldItem s_items[kFoldMaxItems];
  __shared__ FoldTie s_ties[kFoldMaxTies];
  __shared__ double s_x[kFoldMaxItems];
  __shared__ double s_xs, s_E;
  __shared__ int s_pb[2], s_ok;
  const int tid = threadIdx.x;
  if (tid == 0) {
    const unsigned t = atomicAdd(W.look_ctr, 1u);
    s_pb[0] = (int)(t % 3u);
    s_pb[1] = (int)(t / 3u);
  }
  __syncthreads();
  const int p = __builtin_amdgcn_readfirstlane(s_pb[0]), b = __builtin_amdgcn_readfirstlane(s_pb[1]);
  long long *lw = W.look + (size_t)p * A.nblk * kLW;
  long long *mw = lw + (size_t)b * kLW;  // this block's words
  // entry: a phase already marked (the replay hook) skips to the end
  if (tid == 0) s_ok = __hip_atomic_load(W.status + p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0;
  __syncthreads();
  bool ok = s_ok != 0;
  const int base = b * kFoldBlock + tid * kFoldPerThread;
  double d[kFoldPerThread];
  int key[kFoldPerThread];
  long long r[kFoldPerThread];
  int tie_mask = 0, head_mask = 0;
  bool next_starts = true;
  long long carry = 0;
  int item_base = 0, tie_base = 0;
  // the increments, and the block's approximate sum and L1 norm, published at
  // once for the approximate look-back of the blocks after it
  stage_increments<DIR>(A, p, b, sd);
  double tsum = 0.0, tl1 = 0.0;
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    const double v = sd[fold_lds_index(tid * kFoldPerThread + j)];
    tsum += v;
    tl1 += fabs(v);
  }
  double bsum, bl1;
  const double excl = block_excl_sum(tsum, shd, &bsum);
  (void)block_excl_sum(tl1, shd, &bl1);
  if (tid == 0) {
    look_store(mw + kLwAggL, __double_as_longlong(bl1));
    look_store(mw + kLwAggS, __double_as_longlong(bsum));
  }
  // approximate start of the block, and the bound E on the distance between
  // approximate and exact values in it: every value up to the block's end,
  // either way, is at most U = |x0| + the L1 norm so far in magnitude, so
  // each of at most nsteps + 2 kFoldBlock roundings on either path is at most
  // half an ulp of U
  if (tid < 64) {
    double S = A.x0[p], L = fabs(A.x0[p]);
    const bool aok = (b == 0) || approx_look_back(A, lw, p, b, S, L);
    if (tid == 0) {
      look_store(mw + kLwIncL, __double_as_longlong(L + bl1));
      look_store(mw + kLwIncS, __double_as_longlong(S + bsum));
      const double U = (L + bl1) * (1.0 + 0x1p-24);
      int ex = 0;
      frexp(U > 0.0 ? U : 1.0, &ex);
      s_xs = S;
      s_E = ((double)A.nsteps + 2.0 * kFoldBlock) * ldexp(1.0, ex - 52);
      if (!aok) s_ok = 0;
    }
  }
  __syncthreads();
  ok = ok && s_ok != 0;
  const double E = uniform_double(s_E);
  if (ok) {
    // the increments again from LDS (not held across the look-back)
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) d[j] = sd[fold_lds_index(tid * kFoldPerThread + j)];
    double a0 = uniform_double(s_xs) + excl;
    // keys: >= 0 binade grid, -1 exact replay, -2 beyond the end
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) {
      const int s = base + j;
      if (s >= A.nsteps) {
        key[j] = -2;
        r[j] = 0;
        continue;
      }
      const double a1 = a0 + d[j];
      const int k0 = binade_key(a0, E);
      const int k1 = binade_key(a1, E);
      if (k0 >= 0 && k0 == k1) {
        key[j] = k0;
        const int kexp = (k0 >> 1) - 4096;
        const double f = ldexp(d[j], 53 - kexp);
        const double fl = floor(f);
        const double fr = f - fl;
        long long ri = (long long)fl;
        if (fr > 0.5) ri += 1;
        else if (fr == 0.5) tie_mask |= (1 << j);
        r[j] = ri;
      } else {
        key[j] = -1;
        r[j] = 0;
      }
      a0 = a1;
    }
    shi[tid] = key[kFoldPerThread - 1];
    __syncthreads();
    const int prev_key = (tid > 0) ? shi[tid - 1] : -3;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) {
      const int kprev = (j == 0) ? prev_key : key[j - 1];
      const bool h = (key[j] != -2) && (key[j] == -1 || kprev == -1 || kprev == -3 || key[j] != kprev);
      if (h) head_mask |= (1 << j);
    }
    shi[tid] = ((head_mask & 1) || key[0] == -2) ? 1 : 0;
    __syncthreads();
    next_starts = (tid + 1 < kFoldThreads) ? (shi[tid + 1] != 0) : true;
    __syncthreads();
    long long tail = 0;
    int has_head = 0;
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) {
      if (head_mask & (1 << j)) {
        has_head = 1;
        tail = 0;
      }
      tail += r[j];
    }
    const long long seg_incl = block_seg_incl(tail, has_head, shl, shf);
    carry = __shfl_up(seg_incl, 1, 64);
    if ((tid & 63) == 63) shl[kFoldThreads / 64 + (tid >> 6)] = seg_incl;
    __syncthreads();
    if ((tid & 63) == 0) carry = (tid > 0) ? shl[kFoldThreads / 64 + (tid >> 6) - 1] : 0;
    __syncthreads();
    int nheads_tot, nties_tot;
    item_base = block_excl_sum((int)__popc(head_mask), shi, &nheads_tot);
    tie_base = block_excl_sum((int)__popc(tie_mask), shi, &nties_tot);
    if (nheads_tot > kFoldMaxItems || nties_tot > kFoldMaxTies) ok = false;
    if (ok) {
      // the block's item and tie lists, in LDS
      long long P = carry;
      int item = item_base - 1, tcount = tie_base;
#pragma unroll
      for (int j = 0; j < kFoldPerThread; j++) {
        if (key[j] == -2) break;
        const int sl = tid * kFoldPerThread + j;
        if (head_mask & (1 << j)) {
          item++;
          P = 0;
          s_items[item].kind = (key[j] == -1) ? 1 : 0;
          s_items[item].k = (key[j] >= 0) ? (key[j] >> 1) - 4096 : 0;
          s_items[item].sgn = (key[j] >= 0) ? ((key[j] & 1) ? 1 : -1) : 0;
          s_items[item].first = sl;
          s_items[item].tie_first = tcount;
        }
        P += r[j];
        const bool last_of_run = (key[j] >= 0) && ((j + 1 < kFoldPerThread)
                                                       ? (((head_mask >> (j + 1)) & 1) || key[j + 1] == -2)
                                                       : next_starts);
        if (tie_mask & (1 << j)) {
          s_ties[tcount].step = sl;
          s_ties[tcount].P = P;
          s_ties[tcount].up = 0;
          tcount++;
        }
        if (last_of_run) s_items[item].R = P;
      }
    }
    __syncthreads();
    if (tid == 0) {
      s_pb[0] = nheads_tot;
      s_pb[1] = nties_tot;
      // a block that is one run publishes its action at once: the run total
      // plus the ties rounded up, for an even and for an odd start, and its
      // grid key
      const bool pure = ok && nheads_tot == 1 && s_items[0].kind == 0;
      long long a0 = kLookNone, a1 = kLookNone;
      if (pure) {
        long long u0 = 0, u1 = 0;
        for (int t = 0; t < nties_tot; t++) {
          const long long P = s_ties[t].P;
          u0 += (P + u0) & 1;
          u1 += (1 + P + u1) & 1;
        }
        a0 = s_items[0].R + u0;
        a1 = s_items[0].R + u1;
      }
      if (pure) look_store(mw + kLwKey, (long long)((s_items[0].k + 4096) * 2 + (s_items[0].sgn > 0 ? 1 : 0)));
      look_store(mw + kLwA1, a1);
      look_store(mw + kLwA0, a0);
    }
    __syncthreads();
  }
  // wave 0: look back, then the block's own walk; publishes the end value.
  // Lane q holds item q, with its action on an even and an odd start (its
  // ties round to even), so the walk reads uniform values from registers and
  // the end value is published before the ties are resolved for the write.
  if (tid < 64) {
    const int lane = tid, ni = ok ? s_pb[0] : 0, nt = ok ? s_pb[1] : 0;
    int it_kind = 1, it_k = 0, it_sgn = 0, it_t0 = 0, it_t1 = 0;
    long long it_d = 0, it_a0 = 0, it_a1 = 0;
    if (lane < ni) {
      const FoldItem &I = s_items[lane];
      it_kind = I.kind;
      it_k = I.k;
      it_sgn = I.sgn;
      // a single step's increment, from the staged increments (sd holds them
      // until the write)
      it_d = __double_as_longlong(sd[fold_lds_index(I.first)]);
      if (I.kind == 0) {
        it_t0 = I.tie_first;
        it_t1 = (lane + 1 < ni) ? s_items[lane + 1].tie_first : nt;
        long long u0 = 0, u1 = 0;
        for (int t = it_t0; t < it_t1; t++) {
          const long long P = s_ties[t].P;
          u0 += (P + u0) & 1;
          u1 += (1 + P + u1) & 1;
        }
        it_a0 = I.R + u0;
        it_a1 = I.R + u1;
      }
    }
    double x = 0.0;
    bool wok = ok && (b == 0 ? (x = A.x0[p], true) : fold_look_back(A, lw, p, b, x));
    if (W.force_replay == 2 && b == A.nblk / 2) wok = false;  // test hook
    if (wok && b == 0 && lane == 0) A.out[p][A.i0] = x;
    long long my_m = 0;
    double my_x = 0.0;
    for (int q = 0; wok && q < ni; q++) {
      const int kind = __builtin_amdgcn_readlane(it_kind, q);
      if (kind == 1) {
        x = x + __longlong_as_double(lane_read64(it_d, q));
        if (lane == q) my_x = x;
        continue;
      }
      const int k = __builtin_amdgcn_readlane(it_k, q), sgn = __builtin_amdgcn_readlane(it_sgn, q);
      int kx;
      frexp(x, &kx);
      if (kx != k || (x > 0.0 ? 1 : -1) != sgn) {
        wok = false;
        break;
      }
      long long m = (long long)ldexp(x, 53 - k);
      if (lane == q) my_m = m;
      m += (m & 1) ? lane_read64(it_a1, q) : lane_read64(it_a0, q);
      const long long am = m < 0 ? -m : m;
      if (am < (1LL << 52) || am >= (1LL << 53)) {
        wok = false;
        break;
      }
      x = ldexp((double)m, k - 53);
    }
    if (wok && !(x == x)) wok = false;  // a NaN phase takes the serial replay
    if (lane == 0) {
      if (wok) {
        look_store(mw + kLwIncl, __double_as_longlong(x));
      } else {
        __hip_atomic_store(W.status + p, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        look_store(mw + kLwIncl, kLookFail);
      }
      s_ok = wok ? 1 : 0;
    }
    if (wok && lane < ni) {
      // exact start of each run and the rounding of its ties, for the write
      if (it_kind == 1) {
        s_x[lane] = my_x;
      } else {
        s_items[lane].m_start = my_m;
        long long ups = 0;
        for (int t = it_t0; t < it_t1; t++) {
          const int up = (int)((my_m + s_ties[t].P + ups) & 1LL);
          s_ties[t].up = up;
          ups += up;
        }
      }
    }
  }
  __syncthreads();
  if (s_ok) {
    // exact values of every step: runs from their exact starts, single steps
    // from the walk; staged in LDS and stored coalesced
    long long P = carry;
    int item = item_base - 1, tcount = tie_base;
    int c_item = -1, c_k = 0, c_tf = 0;
    long long c_m = 0;
    double outv[kFoldPerThread];
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) {
      outv[j] = 0.0;
      if (key[j] == -2) continue;
      if (head_mask & (1 << j)) {
        item++;
        P = 0;
      }
      P += r[j];
      if (tie_mask & (1 << j)) tcount++;
      if (key[j] >= 0) {
        if (item != c_item) {
          c_item = item;
          c_m = s_items[item].m_start;
          c_k = s_items[item].k;
          c_tf = s_items[item].tie_first;
        }
        long long ups = 0;
        for (int q = c_tf; q < tcount; q++) ups += s_ties[q].up;
        outv[j] = ldexp((double)(c_m + P + ups), c_k - 53);
      } else {
        outv[j] = s_x[item];
      }
    }
#pragma unroll
    for (int j = 0; j < kFoldPerThread; j++) sd[fold_lds_index(tid * kFoldPerThread + j)] = outv[j];
    __syncthreads();
    double *op = A.out[p] + (A.i0 + DIR * (b * kFoldBlock + tid + 1));
#pragma unroll
    for (int q = 0; q < kFoldPerThread; q++) {
      const int s = b * kFoldBlock + tid + q * kFoldThreads;
      if (s < A.nsteps) op[DIR * q * kFoldThreads] = sd[fold_lds_index(tid + q * kFoldThreads)];
    }
  }
}

// serial replay of the phases whose single-pass fold failed
__global__ __launch_bounds__(kFoldThreads) void k_fold_replay(FoldArgs A, FoldWork W) {
  __shared__ double sd[kFoldLds];
  if (W.status[blockIdx.x]) fold_serial_replay(A, W, blockIdx.x, sd, 4096);
}

int fold_exact(const DenseParams &P, const DenseBuffers &B, const FoldWork &W, int nblk_cap, int dir, int nsteps,
               int mode, FoldArgs *defer_replay, hipStream_t s) {
  if (nsteps <= 0) return 0;
  FoldArgs A;
  for (int p = 0; p < 3; p++) {
    A.rate[p] = B.rate[p];
    A.out[p] = B.phase[p];
    A.x0[p] = P.x0[p];
  }
  A.i0 = P.i0;
  A.dir = dir;
  A.nsteps = nsteps;
  A.dt = P.dt;
  A.nblk = (nsteps + kFoldBlock - 1) / kFoldBlock;
  if (A.nblk > nblk_cap) return -1;
  if (mode == 0) {
    if (dir > 0) hipLaunchKernelGGL(k_fold_resolve<1>, dim3(3 * A.nblk), dim3(kFoldThreads), 0, s, A, W);
    else hipLaunchKernelGGL(k_fold_resolve<-1>, dim3(3 * A.nblk), dim3(kFoldThreads), 0, s, A, W);
    // the replay check can ride on the next launch that does not read the
    // phases (the mode-sum setup); otherwise it is a launch of its own
    if (defer_replay) *defer_replay = A;
    else hipLaunchKernelGGL(k_fold_replay, dim3(3), dim3(kFoldThreads), 0, s, A, W);
    return 0;
  }
  if (dir > 0) hipLaunchKernelGGL(k_fold_sums<1>, dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  else hipLaunchKernelGGL(k_fold_sums<-1>, dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_plan, dim3(3), dim3(1024), 0, s, A, W);
  if (dir > 0) hipLaunchKernelGGL((k_fold_local<false, 1>), dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  else hipLaunchKernelGGL((k_fold_local<false, -1>), dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_chain, dim3(3), dim3(kChainThreads), 0, s, A, W);
  if (dir > 0) hipLaunchKernelGGL((k_fold_local<true, 1>), dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  else hipLaunchKernelGGL((k_fold_local<true, -1>), dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  return 0;
}

// per-waveform counters (non-positive Phidot samples, fold fallbacks); a
// kernel rather than a memset so the sequence can be captured in a graph
__global__ void k_clear_flags(int *flags) {
  if (threadIdx.x < 8) flags[threadIdx.x] = 0;
}

void launch_clear_flags(const DenseBuffers &B, hipStream_t s) {
  hipLaunchKernelGGL(k_clear_flags, dim3(1), dim3(64), 0, s, B.flags);
}

// Phases and frequencies per sample (AAKPhase.cc:338-346): t, Phi, Phi+gamma,
// Phi+gamma+alpha, Phidot, Phidot+gammadot, Phidot+gammadot+alphadot, e.
__global__ void k_phase_out(DenseParams P, DenseBuffers B, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = P.N;
  if (i >= N) return;
  const double gim = B.phase[0][i], Phi = B.phase[1][i], alp = B.phase[2][i];
  const double gd = B.rate[0][i], Pd = B.rate[1][i], ad = B.rate[2][i];
  out[i] = sample_time(P, i);
  out[(size_t)N + i] = Phi;
  out[2 * (size_t)N + i] = Phi + gim;
  out[3 * (size_t)N + i] = Phi + gim + alp;
  out[4 * (size_t)N + i] = Pd;
  out[5 * (size_t)N + i] = Pd + gd;
  out[6 * (size_t)N + i] = Pd + gd + ad;
  out[7 * (size_t)N + i] = B.e[i];
}

void launch_phase_out(const DenseParams &P, const DenseBuffers &B, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_phase_out, dim3((P.N + 255) / 256), dim3(256), 0, s, P, B, out);
}

// phases after the plunge buffer stay at their plunge values (AAK.cc:265-269)
__global__ void k_fold_fill(DenseParams P, DenseBuffers B) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + P.last_fold + 2;
  if (i > P.N) return;
  for (int p = 0; p < 3; p++) B.phase[p][i] = B.phase[p][P.ip];
}

void launch_fold_fill(const DenseParams &P, const DenseBuffers &B, hipStream_t s) {
  // only the dense arrays and the phase output read phases past the last
  // folded step; the mode sum of a likelihood stops at the end of the taper
  if (!P.full_dense) return;
  const int n = P.N - (P.last_fold + 2) + 1;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fold_fill, dim3((n + 255) / 256), dim3(256), 0, s, P, B);
}

// ------------------------------------------------------------------ mode sum
// J_{n-2..n+2}(x) by Miller's backward recurrence normalised with
// J_0 + 2 sum J_2k = 1 (A&S 9.12); n = 1 returns (-J_1, J_0, J_1, J_2, J_3).
__device__ inline void bessel_five(int n, double x, double *J) {
  if (x == 0.0) {
    for (int q = 0; q < 5; q++) J[q] = (n - 2 + q == 0) ? 1.0 : 0.0;
    if (n == 1) { J[0] = 0.0; J[1] = 1.0; J[2] = 0.0; J[3] = 0.0; J[4] = 0.0; }
    return;
  }
  const int nmax = n + 2;
  const int nmin = (n >= 2) ? n - 2 : 0;
  int top = nmax + 40 + (int)x;
  top += (top & 1);
  const double inv = 1.0 / x;
  double jp1 = 0.0, j = 1.0e-280, sum = 0.0;
  double w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0;
  for (int k = top; k >= 1; k--) {
    const int q = k - nmin;
    if (q == 0) w0 = j;
    else if (q == 1) w1 = j;
    else if (q == 2) w2 = j;
    else if (q == 3) w3 = j;
    else if (q == 4) w4 = j;
    if ((k & 1) == 0) sum += 2.0 * j;
    const double jm1 = fma((2.0 * k) * inv, j, -jp1);
    jp1 = j;
    j = jm1;
    if (fabs(j) > 1.0e250) {
      j *= 1.0e-250; jp1 *= 1.0e-250; sum *= 1.0e-250;
      w0 *= 1.0e-250; w1 *= 1.0e-250; w2 *= 1.0e-250; w3 *= 1.0e-250; w4 *= 1.0e-250;
    }
  }
  if (nmin == 0) w0 = j;
  sum += j;
  const double norm = 1.0 / sum;
  if (n == 1) {
    J[1] = w0 * norm; J[2] = w1 * norm; J[3] = w2 * norm; J[4] = w3 * norm; J[0] = -J[2];
  } else {
    J[0] = w0 * norm; J[1] = w1 * norm; J[2] = w2 * norm; J[3] = w3 * norm; J[4] = w4 * norm;
  }
}

// Per-sample part of the mode sum (AAK.cc:425-481): everything that does not
// depend on the harmonic, folded into the weights of a_n, b_n, c_n in h_I and
// h_II, so that h = Amp * sum_n [cos(nPhi)(alpha_n W_a + gamma_n W_c) +
// sin(nPhi) beta_n W_b] with a_n = Amp alpha_n cos(nPhi) etc. (AAK.cc:493-504).
struct SampleGeom {
  double e, Phi, nu, gimdot, Amp, ti, cosorbphs;
  double W[6];  // a, b, c weights of h_I, then of h_II
};

__device__ inline double affine_sc(const double *c, double sa, double ca) { return c[0] + c[1] * sa + c[2] * ca; }

// The harmonic-independent quantities need only be accurate (they set
// amplitudes and the Doppler term), except the orientation angle's argument
// 2 (gamma + beta), which is rounded as in the reference; the sample time and
// the dense arrays are read as they are.
__device__ inline void sample_geometry(const DenseParams &P, const DenseBuffers &B, int i, SampleGeom &G) {
#pragma clang fp contract(fast)
  const double ti = sample_time(P, i);
  const double e = B.e[i];
  const double gim = B.phase[0][i], alp = B.phase[2][i];
  G.ti = ti;
  G.e = e;
  G.Phi = B.phase[1][i];
  G.gimdot = B.rate[0][i];
  G.nu = (i == 0) ? B.nu0[0] : B.rate[1][i] * (0.5 / kPi);
  const double coslam = P.coslam;
  const double cosqS = P.cosqS, sinqS = P.sinqS, cosqK = P.cosqK, sinqK = P.sinqK;
  const double halfsqrt3 = P.halfsqrt3;
  double sa, ca;
  sincos(alp, &sa, &ca);
  // orientation of L (AAK.cc:425-433)
  const double cosqL = affine_sc(P.g_cosqL, sa, ca);
  const double sinqL = sqrt(1. - cosqL * cosqL);
  const double phiLup = affine_sc(P.g_phiLup, sa, ca), phiLdown = affine_sc(P.g_phiLdown, sa, ca);
  // cos(phiL - phiS) with phiL = atan2(phiLup, phiLdown) (atan2(0, 0) = 0)
  const double rL2 = phiLup * phiLup + phiLdown * phiLdown;
  const double irL = rL2 > 0. ? rsqrt(rL2) : 0.;
  const double cphiL = rL2 > 0. ? phiLdown * irL : 1., sphiL = phiLup * irL;
  const double Ldotn = cosqL * cosqS + sinqL * sinqS * (cphiL * P.cos_phS + sphiL * P.sin_phS);
  const double Ldotn2 = Ldotn * Ldotn;
  const double betaup = -P.Sdotn + coslam * Ldotn;
  const double betadown = affine_sc(P.g_betadown, sa, ca);
  const double beta = atan2(betaup, betadown);
  // the reference's rounded argument 2 (gamma + beta) of the orientation angle
  const double gb = gim + beta;
  const double gam = 2. * gb;
  double sin2gam, cos2gam;
  sincos(gam, &sin2gam, &cos2gam);
  double cosorbphs = 0.0, FplusI, FcrosI, FplusII, FcrosII, h3;
  if (P.mich) {
    // LISA orbit phase; the shifted angles by angle subtraction from one sincos
    const double orbphs = 2. * kPi * ti / kYear;
    double so, co;
    sincos(orbphs, &so, &co);
    cosorbphs = co * P.cos_phS + so * P.sin_phS;             // cos(orbphs - phiS)
    const double sinorbphs = so * P.cos_phS - co * P.sin_phS;  // sin(orbphs - phiS)
    const double cosorbK = co * P.cos_phK + so * P.sin_phK;    // cos(orbphs - phiK)
    const double cosq = .5 * cosqS - halfsqrt3 * sinqS * cosorbphs;
    // phiw = orbphs + atan2(Yw, Xw): double angles by (X^2 - Y^2, 2XY) / (X^2 + Y^2)
    const double Yw = halfsqrt3 * cosqS + .5 * sinqS * cosorbphs, Xw = sinqS * sinorbphs;
    const double rw = Xw * Xw + Yw * Yw;
    const double irw = rw > 0. ? 1. / rw : 0.;
    const double c2w = rw > 0. ? (Xw * Xw - Yw * Yw) * irw : 1., s2w = 2. * Xw * Yw * irw;
    const double c2o = co * co - so * so, s2o = 2. * so * co;
    const double cos2phi = c2o * c2w - s2o * s2w, sin2phi = s2o * c2w + c2o * s2w;
    const double psiup = .5 * cosqK - halfsqrt3 * sinqK * cosorbK - cosq * P.psi_c0;
    const double psidown = P.psi_c1 - halfsqrt3 * co * P.psi_c2 - halfsqrt3 * so * P.psi_c3;
    const double rp = psiup * psiup + psidown * psidown;
    const double irp = rp > 0. ? 1. / rp : 0.;
    const double cos2psi = rp > 0. ? (psidown * psidown - psiup * psiup) * irp : 1.;
    const double sin2psi = 2. * psidown * psiup * irp;
    const double cosq1 = .5 * (1. + cosq * cosq);
    FplusI = cosq1 * cos2phi * cos2psi - cosq * sin2phi * sin2psi;
    FcrosI = cosq1 * cos2phi * sin2psi + cosq * sin2phi * cos2psi;
    FplusII = cosq1 * sin2phi * cos2psi + cosq * cos2phi * sin2psi;
    FcrosII = cosq1 * sin2phi * sin2psi - cosq * cos2phi * cos2psi;
    h3 = halfsqrt3;
  } else {
    FplusI = P.ldc_c2psi;
    FcrosI = -P.ldc_s2psi;
    FplusII = P.ldc_s2psi;
    FcrosII = P.ldc_c2psi;
    h3 = 1.0;
  }
  G.cosorbphs = cosorbphs;
  G.Amp = B.amp[i];
  // rotation to the wave frame (KSParMap.cc:532-574): cos(rot) = (nxL).(nxS) / N,
  // sin(rot) = (L.(nxS) - S.(nxL)) / N, N = |nxL| |nxS|, all affine or quadratic
  // in (sin alpha, cos alpha)
  const double cn = affine_sc(P.g_rc, sa, ca), sn = affine_sc(P.g_rs, sa, ca);
  const double *q2 = P.g_nxL2;
  const double nxL2 = q2[0] + q2[1] * (sa * sa) + q2[2] * (ca * ca) + q2[3] * sa + q2[4] * ca + q2[5] * (sa * ca);
  const double iN2 = 1. / (nxL2 * P.nxS2);
  const double rot0 = 2. * (cn * cn) * iN2 - 1., rot1 = cn * sn * iN2;
  // h = Ap0 X + Ac0 Y per channel, with Ap0, Ac0 linear in (a, b, c)
  const double XI = h3 * (FplusI * rot0 - FcrosI * rot1), YI = h3 * (FplusI * rot1 + FcrosI * rot0);
  const double XII = h3 * (FplusII * rot0 - FcrosII * rot1), YII = h3 * (FplusII * rot1 + FcrosII * rot0);
  const double pa = -(1. + Ldotn2) * cos2gam, pb = (1. + Ldotn2) * sin2gam, pc = 1. - Ldotn2;
  const double qa = 2. * Ldotn * sin2gam, qb = 2. * Ldotn * cos2gam;
  G.W[0] = pa * XI + qa * YI;
  G.W[1] = pb * XI + qb * YI;
  G.W[2] = pc * XI;
  G.W[3] = pa * XII + qa * YII;
  G.W[4] = pb * XII + qb * YII;
  G.W[5] = pc * XII;
}

// Angle reduced to [-pi, pi] in FP64, returned in FP32 (the reduction error,
// ~k 2.4e-16 rad, is far below FP32 resolution).
__device__ inline float reduce_2pi_f(double x) {
  const double k = rint(x * (1.0 / (2.0 * kPi)));
  return (float)fma(-k, 6.283185307179586, x);
}

// Hardware sine/cosine (v_sin_f32, v_cos_f32; absolute error ~1e-6 for
// |x| <= 4 pi): the response factors they feed set amplitudes, where such
// errors are far below the FP32 harmonic sum's own.
__device__ inline void sincos_hw(float x, float *s, float *c) {
  *s = __sinf(x);
  *c = __cosf(x);
}

// The mixed form's per-sample geometry: the same response as sample_geometry,
// rebuilt in FP32 from unit vectors after the large angles (gamma, alpha, the
// orbital phase) are reduced mod 2 pi in FP64.  With P = cos g E1 + sin g E2,
// Q = L x P (E1, E2 the orbital-plane pair of GeomF32), the harmonic-n
// quadrupole is -a_n (PP - QQ) - b_n (PQ + QP) + c_n (PP + QQ), and a
// channel with (transverse-traceless) response tensor R reads it as R : H.
// For LISA, R = sqrt 3 Lambda(D) with D = (x x - y y)/2 (channel I) or
// (x y + y x)/2 (channel II) on the cartwheel's arm frame (x, y) at orbital
// phase 2 pi t / yr, and Lambda the TT projection about n:
// E Lambda(D) E' = eps D eps' + (eps . eps') (n D n) / 2, eps = E - (n.E) n.
// Without LISA, R = pp - qq and pq + qp on the sky basis.  Nothing here
// divides by a quantity that can vanish, so orientations where the reference's
// angles are ill-conditioned (n along L or S, the source near the detector's
// normal) cost no accuracy.  The phase, its rates, e, t and the amplitude are
// read as they are.
__device__ inline void sample_geometry_mixed(const DenseParams &P, const DenseBuffers &B, int i, SampleGeom &G) {
#pragma clang fp contract(fast)
  const GeomF32 &F = P.gf;
  const double ti = sample_time(P, i);
  G.ti = ti;
  G.e = B.e[i];
  G.Phi = B.phase[1][i];
  G.gimdot = B.rate[0][i];
  G.nu = (i == 0) ? B.nu0[0] : B.rate[1][i] * (0.5 / kPi);
  G.Amp = B.amp[i];
  const float gim = reduce_2pi_f(B.phase[0][i]), alp = reduce_2pi_f(B.phase[2][i]);
  float sa, ca, s2g, c2g;
  sincos_hw(alp, &sa, &ca);
  sincos_hw(2.f * gim, &s2g, &c2g);
  float E1[3], E2[3];
  for (int k = 0; k < 3; k++) {
    E1[k] = F.e1c[k] + F.e1s[k] * sa + F.e1k[k] * ca;
    E2[k] = F.e2s[k] * sa + F.e2k[k] * ca;
  }
  // r[ch][0..2] = E1 R E1, E2 R E2, E1 R E2 of each channel
  float r[2][3];
  float cosorbphs = 0.f;
  if (P.mich) {
    float so, co;
    sincos_hw(reduce_2pi_f(2. * kPi * ti / kYear), &so, &co);
    cosorbphs = co * F.cos_phS + so * F.sin_phS;
    const float h3 = 0.8660254037844386f;
    const float x[3] = {0.5f * so * co, -(co * co + 0.5f * so * so), -h3 * so};
    const float y[3] = {so * so + 0.5f * co * co, -0.5f * so * co, h3 * co};
    auto dot3 = [](const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    const float nx = dot3(F.n, x), ny = dot3(F.n, y), n1 = dot3(F.n, E1), n2 = dot3(F.n, E2);
    const float x1 = dot3(E1, x) - n1 * nx, y1 = dot3(E1, y) - n1 * ny;
    const float x2 = dot3(E2, x) - n2 * nx, y2 = dot3(E2, y) - n2 * ny;
    const float g11 = 1.f - n1 * n1, g22 = 1.f - n2 * n2, g12 = -n1 * n2;
    const float tI = 0.5f * (nx * nx - ny * ny), tII = nx * ny;
    const float s3 = 1.7320508075688772f;
    r[0][0] = s3 * (0.5f * (x1 * x1 - y1 * y1) + 0.5f * g11 * tI);
    r[0][1] = s3 * (0.5f * (x2 * x2 - y2 * y2) + 0.5f * g22 * tI);
    r[0][2] = s3 * (0.5f * (x1 * x2 - y1 * y2) + 0.5f * g12 * tI);
    r[1][0] = s3 * (x1 * y1 + 0.5f * g11 * tII);
    r[1][1] = s3 * (x2 * y2 + 0.5f * g22 * tII);
    r[1][2] = s3 * (0.5f * (x1 * y2 + y1 * x2) + 0.5f * g12 * tII);
  } else {
    const float p1 = F.p[0] * E1[0] + F.p[1] * E1[1] + F.p[2] * E1[2];
    const float q1 = F.q[0] * E1[0] + F.q[1] * E1[1] + F.q[2] * E1[2];
    const float p2 = F.p[0] * E2[0] + F.p[1] * E2[1] + F.p[2] * E2[2];
    const float q2 = F.q[0] * E2[0] + F.q[1] * E2[1] + F.q[2] * E2[2];
    r[0][0] = p1 * p1 - q1 * q1;
    r[0][1] = p2 * p2 - q2 * q2;
    r[0][2] = p1 * p2 - q1 * q2;
    r[1][0] = 2.f * p1 * q1;
    r[1][1] = 2.f * p2 * q2;
    r[1][2] = p1 * q2 + q1 * p2;
  }
  G.cosorbphs = cosorbphs;
  for (int ch = 0; ch < 2; ch++) {
    const float dlt = r[ch][0] - r[ch][1], two12 = 2.f * r[ch][2];
    G.W[3 * ch + 0] = -(c2g * dlt + s2g * two12);
    G.W[3 * ch + 1] = -(c2g * two12 - s2g * dlt);
    G.W[3 * ch + 2] = r[ch][0] + r[ch][1];
  }
}

// Planck taper between plunge and zero time (AAK.cc:535-547)
__device__ inline double taper(const DenseParams &P, int i, double ti) {
  if (i <= P.ip) return 1.0;
  const double tp = P.t_plunge, tz = P.t_zero;
  return exp((tp - tz) / (ti - tp) + (tp - tz) / (ti - tz)) + 1.;
}

// Doppler-shifted harmonic phase exactly as the reference rounds it.
__device__ inline double harmonic_phase(const DenseParams &P, const SampleGeom &G, double nn, double gdpi) {
  if (P.mich) {
    const double fn = nn * G.nu + gdpi;
    const double Doppler = 2. * kPi * fn * kAUsec * P.sinqS * G.cosorbphs;
    return nn * G.Phi + Doppler;
  }
  return nn * G.Phi;
}

// ------------------------------------------------------------------ Taylor form
// Inside one block of samples e varies by a few 1e-5: the amplitude factors
// alpha_n, beta_n, gamma_n are expanded in delta = e - e_ref once per block
// (k_modesum_setup) and evaluated per sample by Horner (k_modesum).  The order
// p is the smallest with remainder bound 7 n (n delta)^{p+1}/(p+1)! <= 1e-17
// (1e-9 for the FP32 harmonic loop of the mixed form).
// Blocks where no order up to kTaylorMax qualifies, or where the Bessel series
// is not used (e > kSeriesMaxE, more than kTaylorModes harmonics), are summed
// per sample by k_modesum_direct.
constexpr int kMsThreads = 256;
constexpr int kTaylorModes = 128;
// Samples per expansion block: 256 << exp_shift (host-set, default 2 in the
// mixed form and 0 in FP64).  The mixed form's looser order bound lets one
// expansion serve 1024 samples (the mean order rises from 2.0 to 2.5 over the
// production prior while the setup work falls fourfold); FP64 keeps 256,
// where wider blocks would send more of them to the per-sample path.
__host__ __device__ inline int exp_shift(const DenseParams &P) { return P.xshift; }
constexpr int kSetupThreads = 64;
// doubles reserved per block in B.ms_coef per harmonic (orders up to kTaylorMax)
constexpr int kCoefStride = 3 * (kTaylorMax + 1);

// Uniform, read-only data through the scalar path: the coefficients of a block
// are the same for all its lanes, so they are fetched with scalar loads into
// SGPRs and feed the FMAs directly.
typedef const __attribute__((address_space(4))) double *cdptr;
typedef const __attribute__((address_space(4))) float *cfptr;
typedef const __attribute__((address_space(4))) int *ciptr;

// B.ms_p[blk]: the Taylor order of 256-sample block blk, or -1 for the
// per-sample path; B.ms_slot[blk]: the slot of its expansion (coefficients
// and e_ref).
__device__ inline int block_order(const DenseBuffers &B, int blk) { return ((ciptr)B.ms_p)[blk]; }
__device__ inline int block_slot(const DenseBuffers &B, int blk) { return ((ciptr)B.ms_slot)[blk]; }

template <int PO, bool MIXED>
__device__ void setup_harmonics(const DenseParams &P, const DenseBuffers &B, int blk, double eref, const double *sq,
                                double *sJ) {
  constexpr int NC = 3 * (PO + 1);
  for (int n = threadIdx.x + 1; n <= P.nmodes; n += kSetupThreads) {
    // series length: the bound for orders n+1, n+2 covers the higher orders
    bessel_run<kSetupThreads>(n, PO, eref, B.bessel_c, B.bessel_lgf, P.bterms[n], sJ);
    const size_t base = (size_t)blk * P.nmodes * kCoefStride;
    // each coefficient is stored as soon as it is formed
    if (MIXED)
      taylor_amplitudes_p<PO, kSetupThreads>(
          n, eref, sq, sJ, CoefStore<PO, float>{reinterpret_cast<float *>(B.ms_coef + base) + (size_t)(n - 1) * NC});
    else
      taylor_amplitudes_p<PO, kSetupThreads>(n, eref, sq, sJ, CoefStore<PO, double>{B.ms_coef + base + (size_t)(n - 1) * NC});
  }
}

// Range of e over samples [i0, i0 + len) below n_end, reduced over the wave
// (every lane holds it); lo > hi when the range holds no signal sample.
__device__ inline void e_range(const DenseBuffers &B, int i0, int len, int n_end, double &lo, double &hi) {
  lo = 1e300;
  hi = -1e300;
  // eight loads in flight per lane before the first use
  constexpr int kU = 8;
  for (int q0 = threadIdx.x; q0 < len; q0 += kU * kSetupThreads) {
    double v[kU];
    bool in[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int q = q0 + u * kSetupThreads, i = i0 + q;
      in[u] = q < len && i < n_end;
      v[u] = in[u] ? B.e[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < kU; u++)
      if (in[u]) {
        lo = fmin(lo, v[u]);
        hi = fmax(hi, v[u]);
      }
  }
  for (int o = 1; o < 64; o <<= 1) {
    lo = fmin(lo, __shfl_xor(lo, o));
    hi = fmax(hi, __shfl_xor(hi, o));
  }
}

// The expansion of one slot at order p about eref (one harmonic per lane).
template <bool MIXED>
__device__ inline void setup_slot(const DenseParams &P, const DenseBuffers &B, int slot, int p, double eref,
                                  double *sJ) {
  double sq[kTaylorMax + 1];
  taylor_sqrt1me2(eref, p, sq);
  switch (p) {
    case 0: setup_harmonics<0, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 1: setup_harmonics<1, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 2: setup_harmonics<2, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 3: setup_harmonics<3, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 4: setup_harmonics<4, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 5: setup_harmonics<5, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 6: setup_harmonics<6, MIXED>(P, B, slot, eref, sq, sJ); break;
    case 7: setup_harmonics<7, MIXED>(P, B, slot, eref, sq, sJ); break;
    default: setup_harmonics<8, MIXED>(P, B, slot, eref, sq, sJ); break;
  }
}

// Block setup, one wave per expansion block of 256 << xs samples: the range
// of e over the block's signal samples, the order p and e_ref (every lane
// derives the same values from a butterfly reduction), then one harmonic per
// lane.  The Bessel run of a lane is staged in LDS (dynamic orders) and the
// expansion is formed in registers.  An expansion block of more than 256
// samples that no order up to kTaylorMax serves (fast evolution near a
// plunge) is listed for k_modesum_split (count in flags[4]); other failures
// (no Bessel series: e > kSeriesMaxE, more than kTaylorModes harmonics) list
// the block's 256-sample blocks for the per-sample path.  A block with no
// signal sample gets p = 0 and no coefficients.
template <bool MIXED>
__global__ __launch_bounds__(kSetupThreads) void k_modesum_setup(DenseParams P, DenseBuffers B, FoldArgs FA,
                                                                 FoldWork FW, int ne) {
  __shared__ double s_J[kTaylorOrders][kSetupThreads];
  const int lane = threadIdx.x, eb = blockIdx.x;
  if (eb >= ne) {
    // workgroups past the expansion blocks: the serial replay of a phase
    // whose single-pass fold failed (the setup reads e only)
    const int p = eb - ne;
    if (FW.status[p]) fold_serial_replay(FA, FW, p, &s_J[0][0], kTaylorOrders * kSetupThreads);
    return;
  }
  const int n_end = min(P.N, P.ip + P.ib);
  const int xs = exp_shift(P), span = kMsThreads << xs, sub = 1 << xs;
  const int nb = (P.N + 2 + kMsThreads - 1) / kMsThreads;
  double lo, hi;
  e_range(B, eb * span, span, n_end, lo, hi);
  const bool any = lo <= hi;
  const bool series = P.bessel_series && P.nmodes <= kTaylorModes && hi <= P.e_series;
  int p = -1;
  double eref = 0.0;
  if (!any) {
    p = 0;
  } else if (series) {
    eref = 0.5 * (lo + hi);
    p = taylor_order(P.nmodes, 0.5 * (hi - lo), MIXED ? kTaylorTolFp32 : kTaylorTolFp64);
  }
  const bool split = p < 0 && series && sub > 1;
  if (lane == 0) {
    B.ms_eref[eb] = eref;
    if (split) B.ms_split[atomicAdd(&B.flags[4], 1)] = eb;
    for (int blk = eb * sub; blk < min(nb, eb * sub + sub); blk++) {
      B.ms_p[blk] = p;
      B.ms_slot[blk] = eb;
      if (p < 0 && !split) B.ms_list[atomicAdd(&B.flags[2], 1)] = blk;
    }
  }
  if (p < 0 || !any) return;
  setup_slot<MIXED>(P, B, eb, p, eref, &s_J[0][lane]);
}

// The 256-sample blocks of the expansion blocks listed by k_modesum_setup,
// each expanded on its own (slot ne + blk), one wave per block over a fixed
// grid; blocks that still need an order above kTaylorMax go to the
// per-sample path.  Returns at once when nothing was listed.
constexpr int kSplitGrid = 256;
template <bool MIXED>
__global__ __launch_bounds__(kSetupThreads) void k_modesum_split(DenseParams P, DenseBuffers B, int ne) {
  const int count = B.flags[4];
  if (count == 0) return;
  __shared__ double s_J[kTaylorOrders][kSetupThreads];
  const int lane = threadIdx.x, sub = 1 << exp_shift(P);
  const int n_end = min(P.N, P.ip + P.ib);
  const int nb = (P.N + 2 + kMsThreads - 1) / kMsThreads;
  for (int k = blockIdx.x; k < count * sub; k += gridDim.x) {
    const int blk = B.ms_split[k / sub] * sub + k % sub;
    if (blk >= nb) continue;
    double lo, hi;
    e_range(B, blk * kMsThreads, kMsThreads, n_end, lo, hi);
    const bool any = lo <= hi;
    int p = 0;
    double eref = 0.0;
    if (any) {
      eref = 0.5 * (lo + hi);
      p = taylor_order(P.nmodes, 0.5 * (hi - lo), MIXED ? kTaylorTolFp32 : kTaylorTolFp64);
    }
    const int slot = ne + blk;
    if (lane == 0) {
      B.ms_p[blk] = p;
      B.ms_slot[blk] = slot;
      B.ms_eref[slot] = eref;
      if (p < 0) B.ms_list[atomicAdd(&B.flags[2], 1)] = blk;
    }
    if (p >= 0 && any) setup_slot<MIXED>(P, B, slot, p, eref, &s_J[0][lane]);
  }
}

int modesum_blocks(const DenseParams &P) { return (P.N + 2 + kMsThreads - 1) / kMsThreads; }
static int expansion_blocks(const DenseParams &P) {
  const int span = kMsThreads << exp_shift(P);
  return (P.N + 2 + span - 1) / span;
}
int modesum_slots(const DenseParams &P) { return expansion_blocks(P) + modesum_blocks(P); }

// Per-sample phase factors shared by both harmonic loops: u = exp(i (Phi + dA))
// with Phi reduced mod 2 pi (Cody-Waite, three parts of 2 pi) and q0 = exp(i dB),
// where the Doppler term of harmonic n is n dA + dB; q_n = q0 u^n by rotation.
struct PhaseSeed {
  double ur, ui, qr, qi, dA, dB;
};

template <bool MIXED>
__device__ inline void phase_seed(const DenseParams &P, const SampleGeom &G, PhaseSeed &S) {
  const double kq = rint(G.Phi * (1.0 / (2.0 * kPi)));
  double phr = fma(-kq, 6.283185307179586, G.Phi);
  phr = fma(-kq, 2.4492935982947064e-16, phr);
  phr = fma(-kq, -5.989539619436679e-33, phr);
  S.dA = 0.0;
  S.dB = 0.0;
  if (P.mich) {
    const double proj = kAUsec * P.sinqS * G.cosorbphs;
    S.dA = 2. * kPi * G.nu * proj;
    S.dB = 2. * G.gimdot * proj;
  }
  if (MIXED) {
    // the seeds of the FP32 loop.  u: sincosf of the FP32 angle, corrected to
    // first order by the angle's FP32 remainder (|remainder| <= 1.2e-7, so
    // the dropped second-order term is below 1e-14).  dB is a small angle (the
    // Doppler shift of the periapsis precession), so FP32 keeps its rotation
    // to ~6e-8.
    const double th = phr + S.dA;
    const float thf = (float)th, tl = (float)(th - (double)thf);
    float su, cu;
    sincosf(thf, &su, &cu);
    S.ui = fmaf(cu, tl, su);
    S.ur = fmaf(-su, tl, cu);
    float qi, qr;
    sincosf((float)S.dB, &qi, &qr);
    S.qi = qi;
    S.qr = qr;
  } else {
    sincos(phr + S.dA, &S.ui, &S.ur);
    sincos(S.dB, &S.qi, &S.qr);
  }
}

// FP64 harmonic loop at Taylor order PO.  The response factors are the same
// for every harmonic, so only the three sums
//   Sa = sum cos(s_n) alpha_n, Sg = sum cos(s_n) gamma_n, Sb = sum sin(s_n) beta_n
// are needed; with alpha_n = sum_j a_nj d^j (and likewise for beta, gamma)
// they are accumulated per power, Sa_j = sum_n cos(s_n) a_nj, one FMA with a
// scalar operand per coefficient, and combined with d^j after the loop.  The
// rotation gives exp(i theta), theta = n Phi + Doppler_n; the reference's
// argument is its rounded s_n = rn(rn(n Phi) + Doppler_n), which differs from
// theta by a few ulp of s_n (~1e-9 rad late in a year-long inspiral).
// E = theta - s_n = fma(n, Phi, -s_n) + Doppler_n, accurate to the rounding of
// Doppler_n (~1e-14 rad), and exp(i s_n) = exp(i theta) (1 - i E) to second
// order in E.
template <int PO>
__device__ inline void harmonics_fp64(int nmodes, cdptr cf, double Phi, double d, const PhaseSeed &S, double *sum) {
  constexpr int NC = 3 * (PO + 1);
  double qr = S.qr, qi = S.qi;
  double Sa[PO + 1], Sb[PO + 1], Sg[PO + 1];
#pragma unroll
  for (int j = 0; j <= PO; j++) Sa[j] = Sb[j] = Sg[j] = 0.0;
  const double ur = S.ur, ui = S.ui;
#pragma unroll 2
  for (int nn = 1; nn <= nmodes; nn++, cf += NC) {
    const double tr = fma(qr, ur, -(qi * ui));
    qi = fma(qr, ui, qi * ur);
    qr = tr;
    const double dn = (double)nn;
    const double D = fma(dn, S.dA, S.dB);
    const double s_ = dn * Phi + D;  // not contracted: rn(rn(n Phi) + D)
    const double E = fma(dn, Phi, -s_) + D;
    const double cnp = fma(E, qi, qr), snp = fma(-E, qr, qi);
#pragma unroll
    for (int j = 0; j <= PO; j++) {
      Sa[j] = fma(cnp, cf[j], Sa[j]);
      Sb[j] = fma(snp, cf[PO + 1 + j], Sb[j]);
      Sg[j] = fma(cnp, cf[2 * (PO + 1) + j], Sg[j]);
    }
  }
  double a = Sa[PO], b = Sb[PO], g = Sg[PO];
#pragma unroll
  for (int j = PO - 1; j >= 0; j--) {
    a = fma(a, d, Sa[j]);
    b = fma(b, d, Sb[j]);
    g = fma(g, d, Sg[j]);
  }
  sum[0] = a;
  sum[1] = b;
  sum[2] = g;
}

// FP32 harmonic loop (mixed precision): rotation and the per-power sums in
// FP32 from FP64-reduced phase factors; no argument correction.
template <int PO>
__device__ inline void harmonics_fp32(int nmodes, cfptr cf, float d, const PhaseSeed &S, double *sum) {
  constexpr int NC = 3 * (PO + 1);
  float qr = (float)S.qr, qi = (float)S.qi;
  float Sa[PO + 1], Sb[PO + 1], Sg[PO + 1];
#pragma unroll
  for (int j = 0; j <= PO; j++) Sa[j] = Sb[j] = Sg[j] = 0.f;
  const float ur = (float)S.ur, ui = (float)S.ui;
#pragma unroll 2
  for (int nn = 1; nn <= nmodes; nn++, cf += NC) {
    const float tr = fmaf(qr, ur, -(qi * ui));
    qi = fmaf(qr, ui, qi * ur);
    qr = tr;
#pragma unroll
    for (int j = 0; j <= PO; j++) {
      Sa[j] = fmaf(qr, cf[j], Sa[j]);
      Sb[j] = fmaf(qi, cf[PO + 1 + j], Sb[j]);
      Sg[j] = fmaf(qr, cf[2 * (PO + 1) + j], Sg[j]);
    }
  }
  float a = Sa[PO], b = Sb[PO], g = Sg[PO];
#pragma unroll
  for (int j = PO - 1; j >= 0; j--) {
    a = fmaf(a, d, Sa[j]);
    b = fmaf(b, d, Sb[j]);
    g = fmaf(g, d, Sg[j]);
  }
  sum[0] = a;
  sum[1] = b;
  sum[2] = g;
}

// Samples past the signal (or the two padding slots) get time and zeros.
__device__ inline bool modesum_tail(const DenseParams &P, const DenseBuffers &B, int i, int n_end) {
  double *hI = B.h, *hII = B.h2;
  if (i >= P.N) {
    B.t[i] = 0.0;
    hI[i] = 0.0;
    hII[i] = 0.0;
    return true;
  }
  if (i >= n_end) {
    B.t[i] = sample_time(P, i);
    hI[i] = 0.0;
    hII[i] = 0.0;
    return true;
  }
  return false;
}

// Taylor-form mode sum for blocks with B.ms_p >= 0.  MIXED: the harmonic loop
// runs in FP32 (phases, their reduction mod 2 pi, the block expansions and the
// final scaling stay FP64).
template <bool MIXED>
__device__ inline void modesum_block(const DenseParams &P, const DenseBuffers &B, int blk) {
  const int p = block_order(B, blk);
  if (p < 0) return;  // k_modesum_direct sums this block
  const int slot = block_slot(B, blk);
  const int i = blk * kMsThreads + threadIdx.x;
  if (i >= P.N + 2) return;
  const int n_end = min(P.N, P.ip + P.ib);  // samples [0, n_end) carry the signal
  if (modesum_tail(P, B, i, n_end)) return;
  SampleGeom G;
  if (MIXED) sample_geometry_mixed(P, B, i, G);
  else sample_geometry(P, B, i, G);
  B.t[i] = G.ti;
  PhaseSeed S;
  phase_seed<MIXED>(P, G, S);
  const double d = G.e - B.ms_eref[slot];
  const size_t base = (size_t)slot * P.nmodes * kCoefStride;
  double sum[3];
  if (MIXED) {
    cfptr cf = (cfptr)(B.ms_coef + base);
    const float fd = (float)d;
    switch (p) {
      case 0: harmonics_fp32<0>(P.nmodes, cf, fd, S, sum); break;
      case 1: harmonics_fp32<1>(P.nmodes, cf, fd, S, sum); break;
      case 2: harmonics_fp32<2>(P.nmodes, cf, fd, S, sum); break;
      case 3: harmonics_fp32<3>(P.nmodes, cf, fd, S, sum); break;
      case 4: harmonics_fp32<4>(P.nmodes, cf, fd, S, sum); break;
      case 5: harmonics_fp32<5>(P.nmodes, cf, fd, S, sum); break;
      case 6: harmonics_fp32<6>(P.nmodes, cf, fd, S, sum); break;
      case 7: harmonics_fp32<7>(P.nmodes, cf, fd, S, sum); break;
      default: harmonics_fp32<8>(P.nmodes, cf, fd, S, sum); break;
    }
  } else {
    cdptr cf = (cdptr)(B.ms_coef + base);
    switch (p) {
      case 0: harmonics_fp64<0>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 1: harmonics_fp64<1>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 2: harmonics_fp64<2>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 3: harmonics_fp64<3>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 4: harmonics_fp64<4>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 5: harmonics_fp64<5>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 6: harmonics_fp64<6>(P.nmodes, cf, G.Phi, d, S, sum); break;
      case 7: harmonics_fp64<7>(P.nmodes, cf, G.Phi, d, S, sum); break;
      default: harmonics_fp64<8>(P.nmodes, cf, G.Phi, d, S, sum); break;
    }
  }
  const double accI = G.W[0] * sum[0] + G.W[1] * sum[1] + G.W[2] * sum[2];
  const double accII = G.W[3] * sum[0] + G.W[4] * sum[1] + G.W[5] * sum[2];
  const double sc = (i <= P.ip) ? G.Amp : G.Amp / taper(P, i, G.ti);
  B.h[i] = sc * accI;
  B.h2[i] = sc * accII;
}

// One workgroup per 256-sample block, XCD-aware: workgroups are dispatched in
// index order and dealt to the 8 XCDs round-robin (workgroup g runs on XCD
// g % 8), so g is mapped to the block that puts the 2^xs sample blocks of one
// expansion block on one XCD (their Taylor coefficients are fetched into one
// L2) while the blocks in flight at any moment stay one contiguous stretch of
// the waveform, spread over every XCD (no XCD idles past a plunge).  Blocks
// beyond the last whole group of 8 << xs keep their index.
template <bool MIXED>
__global__ __launch_bounds__(kMsThreads) void k_modesum(DenseParams P, DenseBuffers B, int nb) {
  const int xs = exp_shift(P), g = blockIdx.x;
  const int full = nb / (8 << xs) * (8 << xs);
  int blk = g;
  if (g < full) {
    const int x = g & 7, r = g >> 3;
    blk = ((((r >> xs) << 3) | x) << xs) | (r & ((1 << xs) - 1));
  }
  modesum_block<MIXED>(P, B, blk);
}

// Per-sample mode sum for the 256-sample blocks listed by k_modesum_setup (B.ms_p < 0):
// eight lanes per sample, each summing every eighth harmonic (the per-harmonic
// Bessel and phase evaluation is the long part), partial sums combined by
// shuffles.  A fixed grid walks the list, 32 samples per workgroup.
constexpr int kDirectGroups = 8;
constexpr int kDirectSamples = kMsThreads / kDirectGroups;
constexpr int kDirectGrid = 1024;

__device__ inline void modesum_direct_part(const DenseParams &P, const DenseBuffers &B, const SampleGeom &G, int g,
                                           double *hI, double *hII) {
  const double e = G.e, sq1me2 = sqrt(1 - e * e), gdpi = G.gimdot / kPi;
  double accI = 0.0, accII = 0.0, J[5];
  for (int nn = 1 + g; nn <= P.nmodes; nn += kDirectGroups) {
    const double nPhi = harmonic_phase(P, G, (double)nn, gdpi);
    if (P.bessel_series && e <= P.e_series)
      bessel_five_series(nn, e, B.bessel_c, B.bessel_lgf, P.bterms[nn], P.bterms[0], J);
    else
      bessel_five(nn, nn * e, J);
    double snp, cnp;
    sincos(nPhi, &snp, &cnp);
    const double al = -nn * (J[0] - 2. * e * J[1] + 2. / nn * J[2] + 2. * e * J[3] - J[4]);
    const double be = -nn * sq1me2 * (J[0] - 2. * J[2] + J[4]);
    const double ga = 2. * J[2];
    accI += cnp * (al * G.W[0] + ga * G.W[2]) + snp * (be * G.W[1]);
    accII += cnp * (al * G.W[3] + ga * G.W[5]) + snp * (be * G.W[4]);
  }
  *hI = accI;
  *hII = accII;
}

__global__ __launch_bounds__(kMsThreads) void k_modesum_direct(DenseParams P, DenseBuffers B) {
  const int count = B.flags[2];
  const int span = kMsThreads;           // samples per listed block
  const int per = span / kDirectSamples;  // workgroup passes per listed block
  const int units = count * per;
  const int g = threadIdx.x % kDirectGroups, sl = threadIdx.x / kDirectGroups;
  const int n_end = min(P.N, P.ip + P.ib);
  for (int u = blockIdx.x; u < units; u += gridDim.x) {
    const int blk = B.ms_list[u / per];
    const int i = blk * span + (u % per) * kDirectSamples + sl;
    // every lane takes part in the shuffles below; inactive ones add zeros
    bool work = i < P.N + 2;
    if (work && g == 0) work = !modesum_tail(P, B, i, n_end);
    else if (work) work = (i < P.N && i < n_end);
    double accI = 0.0, accII = 0.0;
    SampleGeom G;
    if (work) {
      sample_geometry(P, B, i, G);
      modesum_direct_part(P, B, G, g, &accI, &accII);
    }
#pragma unroll
    for (int o = 1; o < kDirectGroups; o <<= 1) {
      accI += __shfl_xor(accI, o, 64);
      accII += __shfl_xor(accII, o, 64);
    }
    if (work && g == 0) {
      B.t[i] = G.ti;
      const double sc = (i <= P.ip) ? G.Amp : G.Amp / taper(P, i, G.ti);
      B.h[i] = sc * accI;
      B.h2[i] = sc * accII;
    }
  }
}

// Workgroups of k_modesum the device runs at once (CUs x occupancy), per precision.
void launch_modesum(const DenseParams &P, const DenseBuffers &B, const FoldArgs *replay, const FoldWork *fw,
                    hipStream_t s) {
  const int nb = modesum_blocks(P), ne = expansion_blocks(P);
  const FoldArgs FA = replay ? *replay : FoldArgs{};
  const FoldWork FW = fw ? *fw : FoldWork{};
  const int ng = ne + (replay ? 3 : 0);
  if (P.mixed) {
    hipLaunchKernelGGL(k_modesum_setup<true>, dim3(ng), dim3(kSetupThreads), 0, s, P, B, FA, FW, ne);
    if (exp_shift(P) > 0) hipLaunchKernelGGL(k_modesum_split<true>, dim3(kSplitGrid), dim3(kSetupThreads), 0, s, P, B, ne);
    hipLaunchKernelGGL(k_modesum<true>, dim3(nb), dim3(kMsThreads), 0, s, P, B, nb);
  } else {
    hipLaunchKernelGGL(k_modesum_setup<false>, dim3(ng), dim3(kSetupThreads), 0, s, P, B, FA, FW, ne);
    if (exp_shift(P) > 0)
      hipLaunchKernelGGL(k_modesum_split<false>, dim3(kSplitGrid), dim3(kSetupThreads), 0, s, P, B, ne);
    hipLaunchKernelGGL(k_modesum<false>, dim3(nb), dim3(kMsThreads), 0, s, P, B, nb);
  }
  hipLaunchKernelGGL(k_modesum_direct, dim3(kDirectGrid), dim3(kMsThreads), 0, s, P, B);
}

// ------------------------------------------------------------------ likelihood
// <h w, d> and <h w, h w> over two channels; the weighted spectrum is written
// back (the reference leaves h~ w in its FFT buffer).
constexpr int kDotThreads = 256;
constexpr int kDotBlocks = 1024;

__global__ __launch_bounds__(kDotThreads) void k_weighted_dot(const double2 *hf, const double2 *data, const double *w,
                                                              int nbins, double2 *hw_out, double *partials) {
  __shared__ double s_dh[kDotThreads], s_hh[kDotThreads];
  double dh = 0.0, hh = 0.0;
  const int total = 2 * nbins;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < total; q += gridDim.x * blockDim.x) {
    const double2 h = hf[q];
    const double ww = w[q];
    const double2 x = make_double2(h.x * ww, h.y * ww);
    const double2 dd = data[q];
    dh += x.x * dd.x + x.y * dd.y;
    hh += x.x * x.x + x.y * x.y;
    if (hw_out) hw_out[q] = x;
  }
  s_dh[threadIdx.x] = dh;
  s_hh[threadIdx.x] = hh;
  __syncthreads();
  for (int o = kDotThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_dh[threadIdx.x] += s_dh[threadIdx.x + o];
      s_hh[threadIdx.x] += s_hh[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = s_dh[0];
    partials[2 * blockIdx.x + 1] = s_hh[0];
  }
}

// out = 4 x (sum of the partial pairs); NaN when the waveform was rejected
// on the device (flags[3], a non-positive radial rate at t = 0)
// flags: the waveform's flags (flags[3]: rejected on the device, the result
// is NaN) or null; clear: flags to zero once read (the lane's, so that its
// next waveform starts from zeros without a clearing launch) or null.
__global__ __launch_bounds__(kDotThreads) void k_dot_final(const double *partials, int nparts, double *out,
                                                           const int *flags, int *clear) {
  __shared__ double s_dh[kDotThreads], s_hh[kDotThreads];
  double dh = 0.0, hh = 0.0;
  // four pairs in flight per lane before the first add; the adds keep the
  // order of one pair per step
  constexpr int kU = 4;
  for (int q0 = threadIdx.x; q0 < nparts; q0 += kU * kDotThreads) {
    double2 v[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const int q = q0 + u * kDotThreads;
      v[u] = (q < nparts) ? reinterpret_cast<const double2 *>(partials)[q] : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int u = 0; u < kU; u++)
      if (q0 + u * kDotThreads < nparts) {
        dh += v[u].x;
        hh += v[u].y;
      }
  }
  s_dh[threadIdx.x] = dh;
  s_hh[threadIdx.x] = hh;
  __syncthreads();
  for (int o = kDotThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s_dh[threadIdx.x] += s_dh[threadIdx.x + o];
      s_hh[threadIdx.x] += s_hh[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const bool bad = flags && flags[3] != 0;
    out[0] = bad ? __builtin_nan("") : 4.0 * s_dh[0];
    out[1] = bad ? __builtin_nan("") : 4.0 * s_hh[0];
  }
  __syncthreads();
  if (clear && threadIdx.x < 8) clear[threadIdx.x] = 0;
}

void launch_dot_final(const double *partials, int nparts, double *out, const int *flags, int *clear, hipStream_t s) {
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotThreads), 0, s, partials, nparts, out, flags, clear);
}

void launch_weighted_dot(const double2 *hf, const double2 *data, const double *w, int nbins, double2 *hw_out,
                         double *partials, double *out, const int *flags, int *clear, hipStream_t s) {
  const int need = (2 * nbins + kDotThreads - 1) / kDotThreads;
  const int nb = need < kDotBlocks ? need : kDotBlocks;
  hipLaunchKernelGGL(k_weighted_dot, dim3(nb), dim3(kDotThreads), 0, s, hf, data, w, nbins, hw_out, partials);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotThreads), 0, s, partials, nb, out, flags, clear);
}

}  // namespace aak
