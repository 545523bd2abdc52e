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
//   k_modesum      per-sample harmonic sum with Bessel amplitudes, Doppler
//                  phases, LISA response and Planck taper (AAK.cc:409-547)
//   k_weighted_dot PSD-weighted inner products of the rFFT output
//                  (AAK_manager.cu:352-446 semantics)
//
// Compiled with -ffp-contract=off: every expression that feeds the phases
// must round exactly like the reference CPU build.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "aak_dense.hpp"
#include "aak_math.hpp"

namespace aak {

// ------------------------------------------------------------------ rates
constexpr int kRatesThreads = 256;
constexpr int kStageCap = 320;  // knot intervals staged in LDS per block

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
  const int i_lo = blockIdx.x * kRatesThreads;
  if (i_lo > P.N) return;
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
  if (staged) {
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
  B.v[i] = v;
  B.M[i] = M;
  B.S[i] = S;
  const double Y = P.coslam;
  const double T = PN::dtdm(v, e, Y, S);
  const double R = PN::drdm(v, e, Y, S);
  const double Th = PN::dthetadm(v, e, Y, S);
  const double Ph = PN::dphidm(v, e, Y, S);
  const double phidot = R / T / M;
  B.rate[0][i] = (Th - R) / T / M;
  B.rate[1][i] = phidot;
  B.rate[2][i] = (Ph - Th) / T / M;
  if (i > P.i0 && i <= P.guard_last && !(phidot > 0.)) atomicAdd(&B.flags[0], 1);
}

void launch_rates(const DenseParams &P, const DenseBuffers &B, hipStream_t s) {
  const int nb = (P.N + 1 + kRatesThreads - 1) / kRatesThreads;
  hipLaunchKernelGGL(k_rates, dim3(nb), dim3(kRatesThreads), 0, s, P, B);
}

// Non-positive radial rates take the previous (already corrected) value
// (AAK.cc:254): sample i in (i0, guard_last] with Phidot <= 0 (or NaN) gets the
// rate of the last sample j in (i0, i] with Phidot > 0, or of i0 if there is
// none.  That source index is a running maximum, computed as a two-pass
// parallel scan; both passes return at once when k_rates counted no such
// sample, so they are always enqueued and need no host round trip.
constexpr int kGuardThreads = 256;
constexpr int kGuardPer = 16;
constexpr int kGuardBlock = kGuardThreads * kGuardPer;

__device__ inline bool phidot_ok(double r) { return r > 0.; }

__global__ __launch_bounds__(kGuardThreads) void k_guard_last(DenseParams P, DenseBuffers B) {
  if (B.flags[0] == 0) return;
  __shared__ int s_m[kGuardThreads];
  const double *r = B.rate[1];
  const int base = P.i0 + 1 + blockIdx.x * kGuardBlock + threadIdx.x * kGuardPer;
  int last = -1;
  for (int q = 0; q < kGuardPer; q++) {
    const int i = base + q;
    if (i <= P.guard_last && phidot_ok(r[i])) last = i;
  }
  s_m[threadIdx.x] = last;
  __syncthreads();
  for (int o = kGuardThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = max(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) B.guard_blk[blockIdx.x] = s_m[0];
}

__global__ __launch_bounds__(kGuardThreads) void k_guard_fix(DenseParams P, DenseBuffers B) {
  if (B.flags[0] == 0) return;
  __shared__ int s_m[kGuardThreads];
  __shared__ int s_carry;
  double *r = B.rate[1];
  // last good index before this block
  int c = P.i0;
  for (int b = threadIdx.x; b < (int)blockIdx.x; b += kGuardThreads) c = max(c, B.guard_blk[b]);
  s_m[threadIdx.x] = c;
  __syncthreads();
  for (int o = kGuardThreads / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) s_m[threadIdx.x] = max(s_m[threadIdx.x], s_m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) s_carry = s_m[0];
  __syncthreads();
  const int carry = s_carry;
  __syncthreads();
  const int base = P.i0 + 1 + blockIdx.x * kGuardBlock + threadIdx.x * kGuardPer;
  double v[kGuardPer];
  int last = -1;
  for (int q = 0; q < kGuardPer; q++) {
    const int i = base + q;
    v[q] = (i <= P.guard_last) ? r[i] : 1.0;
    if (i <= P.guard_last && phidot_ok(v[q])) last = i;
  }
  // inclusive max-scan of the per-thread last good index
  s_m[threadIdx.x] = last;
  __syncthreads();
  for (int o = 1; o < kGuardThreads; o <<= 1) {
    const int add = (threadIdx.x >= o) ? s_m[threadIdx.x - o] : -1;
    __syncthreads();
    s_m[threadIdx.x] = max(s_m[threadIdx.x], add);
    __syncthreads();
  }
  int src = max(carry, threadIdx.x > 0 ? s_m[threadIdx.x - 1] : -1);
  for (int q = 0; q < kGuardPer; q++) {
    const int i = base + q;
    if (i > P.guard_last) break;
    if (phidot_ok(v[q])) {
      src = i;
    } else {
      r[i] = r[src];
    }
  }
}

int guard_blocks(const DenseParams &P) {
  const int n = P.guard_last - P.i0;
  return n > 0 ? (n + kGuardBlock - 1) / kGuardBlock : 0;
}

void launch_phidot_guard(const DenseParams &P, const DenseBuffers &B, hipStream_t s) {
  const int nb = guard_blocks(P);
  if (nb == 0) return;
  hipLaunchKernelGGL(k_guard_last, dim3(nb), dim3(kGuardThreads), 0, s, P, B);
  hipLaunchKernelGGL(k_guard_fix, dim3(nb), dim3(kGuardThreads), 0, s, P, B);
}

// ------------------------------------------------------------------ folds
struct FoldArgs {
  const double *rate[3];
  double *out[3];
  double x0[3];
  int i0, dir, nsteps, nblk;
  double dt;
};

__device__ inline double fold_incr(const double *rate, int i0, int dir, int s, double dt) {
  const int i = i0 + dir * s;
  const double r = rate[i];
  const double rm = (s == 0) ? r : rate[i - dir];
  const double d = (1.5 * r - .5 * rm) * dt;
  return dir > 0 ? d : -d;
}

__device__ inline int out_index(const FoldArgs &A, int s) { return A.i0 + A.dir * (s + 1); }

// block-wide scans over kFoldThreads lanes through LDS (Hillis-Steele)
template <class T>
__device__ inline T block_excl_sum(T v, T *sh, T *total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = 1; off < kFoldThreads; off <<= 1) {
    const T add = (tid >= off) ? sh[tid - off] : T(0);
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  const T incl = sh[tid];
  if (total) *total = sh[kFoldThreads - 1];
  __syncthreads();
  return incl - v;
}

// approximate block sums and prefix ranges
__global__ __launch_bounds__(kFoldThreads) void k_fold_sums(FoldArgs A, FoldWork W) {
  __shared__ double sh[kFoldThreads];
  const int p = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  const int base = b * kFoldBlock + tid * kFoldPerThread;
  double run = 0.0, mn = 0.0, mx = 0.0;
  for (int j = 0; j < kFoldPerThread; j++) {
    const int s = base + j;
    if (s < A.nsteps) {
      run += fold_incr(A.rate[p], A.i0, A.dir, s, A.dt);
      mn = fmin(mn, run);
      mx = fmax(mx, run);
    }
  }
  double tot;
  const double off = block_excl_sum(run, sh, &tot);
  double lmn = off + mn, lmx = off + mx;
  sh[tid] = lmn;
  __syncthreads();
  for (int o = kFoldThreads / 2; o > 0; o >>= 1) {
    if (tid < o) sh[tid] = fmin(sh[tid], sh[tid + o]);
    __syncthreads();
  }
  const double bmn = sh[0];
  __syncthreads();
  sh[tid] = lmx;
  __syncthreads();
  for (int o = kFoldThreads / 2; o > 0; o >>= 1) {
    if (tid < o) sh[tid] = fmax(sh[tid], sh[tid + o]);
    __syncthreads();
  }
  if (tid == 0) {
    W.bsum[p * A.nblk + b] = tot;
    W.bmin[p * A.nblk + b] = bmn;
    W.bmax[p * A.nblk + b] = sh[0];
  }
}

// approximate block start values and the error bound E
__global__ __launch_bounds__(1024) void k_fold_plan(FoldArgs A, FoldWork W) {
  __shared__ double sh[1024];
  __shared__ double shm[1024];
  const int p = blockIdx.x, tid = threadIdx.x;
  const int per = (A.nblk + 1023) / 1024;
  double loc = 0.0;
  for (int j = 0; j < per; j++) {
    const int b = tid * per + j;
    if (b < A.nblk) loc += W.bsum[p * A.nblk + b];
  }
  sh[tid] = loc;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const double add = (tid >= off) ? sh[tid - off] : 0.0;
    __syncthreads();
    sh[tid] += add;
    __syncthreads();
  }
  double run = A.x0[p] + (sh[tid] - loc);
  double amax = fabs(A.x0[p]);
  for (int j = 0; j < per; j++) {
    const int b = tid * per + j;
    if (b < A.nblk) {
      W.xstart[p * A.nblk + b] = run;
      amax = fmax(amax, fmax(fabs(run + W.bmin[p * A.nblk + b]), fabs(run + W.bmax[p * A.nblk + b])));
      run += W.bsum[p * A.nblk + b];
    }
  }
  shm[tid] = amax;
  __syncthreads();
  for (int o = 512; o > 0; o >>= 1) {
    if (tid < o) shm[tid] = fmax(shm[tid], shm[tid + o]);
    __syncthreads();
  }
  if (tid == 0) {
    int ex = 0;
    frexp(shm[0] > 0.0 ? shm[0] : 1.0, &ex);
    W.E[p] = ((double)A.nsteps + 2.0 * kFoldBlock) * ldexp(1.0, ex - 52);
    W.status[p] = 0;
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

// Classify steps, build the block's item list (PLAN) or write the exact
// fold values of run steps (WRITE, after the chain resolved the run starts).
template <bool WRITE>
__global__ __launch_bounds__(kFoldThreads) void k_fold_local(FoldArgs A, FoldWork W) {
  __shared__ double shd[kFoldThreads];
  __shared__ long long shl[kFoldThreads];
  __shared__ int shi[kFoldThreads];
  __shared__ int shf[kFoldThreads];
  const int p = blockIdx.y, b = blockIdx.x, tid = threadIdx.x;
  if (W.status[p]) return;
  const int base = b * kFoldBlock + tid * kFoldPerThread;
  const double E = W.E[p];
  double d[kFoldPerThread];
  double tsum = 0.0;
#pragma unroll
  for (int j = 0; j < kFoldPerThread; j++) {
    const int s = base + j;
    d[j] = (s < A.nsteps) ? fold_incr(A.rate[p], A.i0, A.dir, s, A.dt) : 0.0;
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
  shl[tid] = tail;
  shf[tid] = has_head;
  __syncthreads();
  for (int off = 1; off < kFoldThreads; off <<= 1) {
    long long v = 0;
    int f = 0;
    if (tid >= off) {
      v = shl[tid - off];
      f = shf[tid - off];
    }
    __syncthreads();
    if (tid >= off && !shf[tid]) {
      shl[tid] += v;
      shf[tid] |= f;
    }
    __syncthreads();
  }
  const long long carry = (tid > 0) ? shl[tid - 1] : 0;
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
      A.out[p][out_index(A, base + j)] = ldexp((double)m, c_k - 53);
    }
  }
}

// Exact start of every run, resolution of the exact ties, FP64 replay of the
// single steps near binade boundaries.  The item and tie lists of all blocks
// are gathered into LDS; consecutive runs on one binade grid without ties
// form a segment whose total is a segmented int64 scan, so the one serial
// lane only visits segment heads (binade changes, ties, single steps); the
// run starts inside a segment are filled in parallel and re-checked against
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

// exclusive scan of v[0..n) in place (n <= kChainMaxBlocks), returns the total
__device__ inline int block_exclusive_scan(int *v, int n, int *s_tmp) {
  const int tid = threadIdx.x;
  const int per = (n + kChainThreads - 1) / kChainThreads;
  const int lo = min(n, tid * per), hi = min(n, lo + per);
  int sum = 0;
  for (int j = lo; j < hi; j++) sum += v[j];
  s_tmp[tid] = sum;
  __syncthreads();
  for (int o = 1; o < kChainThreads; o <<= 1) {
    const int add = (tid >= o) ? s_tmp[tid - o] : 0;
    __syncthreads();
    s_tmp[tid] += add;
    __syncthreads();
  }
  int run = s_tmp[tid] - sum;
  for (int j = lo; j < hi; j++) {
    const int c = v[j];
    v[j] = run;
    run += c;
  }
  const int total = s_tmp[kChainThreads - 1];
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kChainThreads) void k_fold_chain(FoldArgs A, FoldWork W) {
  __shared__ int s_io[kChainMaxBlocks + 1], s_to[kChainMaxBlocks + 1];
  __shared__ int s_tmp[kChainThreads];
  __shared__ long long s_ltmp[kChainThreads];
  __shared__ int s_ftmp[kChainThreads];
  __shared__ FoldItem s_items[kChainMaxItems];
  __shared__ FoldTie s_ties[kChainMaxTies];
  __shared__ long long s_pref[kChainMaxItems];  // exclusive R prefix inside the segment
  __shared__ int s_head[kChainMaxItems + 1];     // head flags, then the compacted head list
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
    TI = block_exclusive_scan(s_io, nblk, s_tmp);
    TT = block_exclusive_scan(s_to, nblk, s_tmp);
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
  // gather: items get their block and a block-global tie offset
  for (int b = tid; b < nblk; b += kChainThreads) {
    const int io = s_io[b], ni = s_io[b + 1] - io, to = s_to[b];
    for (int j = 0; j < ni; j++) {
      FoldItem it = gitems[(size_t)b * kFoldMaxItems + j];
      it.blk = b;
      it.tie_first += to;
      s_items[io + j] = it;
    }
    const int nt = s_to[b + 1] - to;
    for (int j = 0; j < nt; j++) s_ties[to + j] = gties[(size_t)b * kFoldMaxTies + j];
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
  // segmented exclusive scan of R over items (per-thread chunks, then lanes)
  {
    const int per = (TI + kChainThreads - 1) / kChainThreads;
    const int lo = min(TI, tid * per), hi = min(TI, lo + per);
    long long tail = 0;
    int has = 0;
    for (int j = lo; j < hi; j++) {
      if (s_head[j]) {
        has = 1;
        tail = 0;
      }
      tail += s_items[j].kind == 0 ? s_items[j].R : 0;
    }
    s_ltmp[tid] = tail;
    s_ftmp[tid] = has;
    __syncthreads();
    for (int off = 1; off < kChainThreads; off <<= 1) {
      long long v = 0;
      int f = 0;
      if (tid >= off) {
        v = s_ltmp[tid - off];
        f = s_ftmp[tid - off];
      }
      __syncthreads();
      if (tid >= off && !s_ftmp[tid]) {
        s_ltmp[tid] += v;
        s_ftmp[tid] |= f;
      }
      __syncthreads();
    }
    long long run = (tid > 0) ? s_ltmp[tid - 1] : 0;
    for (int j = lo; j < hi; j++) {
      if (s_head[j]) run = 0;
      s_pref[j] = run;
      run += s_items[j].kind == 0 ? s_items[j].R : 0;
    }
  }
  __syncthreads();
  // compact the head list (flags -> indices)
  {
    const int per = (TI + kChainThreads - 1) / kChainThreads;
    const int lo = min(TI, tid * per), hi = min(TI, lo + per);
    int cnt = 0;
    for (int j = lo; j < hi; j++) cnt += s_head[j];
    s_tmp[tid] = cnt;
    __syncthreads();
    for (int o = 1; o < kChainThreads; o <<= 1) {
      const int add = (tid >= o) ? s_tmp[tid - o] : 0;
      __syncthreads();
      s_tmp[tid] += add;
      __syncthreads();
    }
    int pos = s_tmp[tid] - cnt;
    int flags[8];
    const int nloc = hi - lo;
    for (int j = 0; j < nloc && j < 8; j++) flags[j] = s_head[lo + j];
    __syncthreads();
    for (int j = 0; j < nloc; j++) {
      const int f = (j < 8) ? flags[j] : 0;
      if (f) s_head[pos++] = lo + j;
    }
    if (tid == kChainThreads - 1) s_nheads = s_tmp[kChainThreads - 1];
  }
  __syncthreads();
  // serial walk over the segment heads
  if (tid == 0) {
    const int nh = s_nheads;
    double x = A.x0[p];
    A.out[p][A.i0] = x;
    int ok = 1;
    for (int q = 0; q < nh && ok; q++) {
      const int h = s_head[q];
      const int last = ((q + 1 < nh) ? s_head[q + 1] : TI) - 1;
      FoldItem &I = s_items[h];
      // the segment's later runs add their totals to the head's end value
      // (s_pref is the exclusive prefix from the head, so it includes R_h)
      const long long extra = (last > h) ? s_pref[last] + s_items[last].R - I.R : 0;
      ok = chain_item<FoldItem &>(A, p, I.blk, I, s_ties + I.tie_first, nties_of(h), x, extra) ? 1 : 0;
    }
    s_ok = ok;
    if (!ok) W.status[p] = 1;
  }
  __syncthreads();
  if (!s_ok) return;
  // run starts inside segments, re-checked against their binade
  const int nh = s_nheads;
  int bad = 0;
  for (int q = tid; q < nh; q += kChainThreads) {
    const int h = s_head[q];
    const int last = ((q + 1 < nh) ? s_head[q + 1] : TI) - 1;
    const long long mh = s_items[h].m_start;
    for (int i = h + 1; i <= last; i++) {
      const long long m = mh + s_pref[i];
      const long long am = m < 0 ? -m : m;
      if (am < (1LL << 52) || am >= (1LL << 53)) bad = 1;
      s_items[i].m_start = m;
    }
  }
  if (bad) W.status[p] = 1;
  __syncthreads();
  for (int b = tid; b < nblk; b += kChainThreads) {
    const int io = s_io[b], ni = s_io[b + 1] - io;
    for (int j = 0; j < ni; j++) gitems[(size_t)b * kFoldMaxItems + j].m_start = s_items[io + j].m_start;
    const int to = s_to[b], nt = s_to[b + 1] - to;
    for (int j = 0; j < nt; j++) gties[(size_t)b * kFoldMaxTies + j].up = s_ties[to + j].up;
  }
}

// Fallback: plain serial replay of the whole fold (used when the item
// bookkeeping overflows; correct for any input).
__global__ __launch_bounds__(256) void k_fold_serial(FoldArgs A, FoldWork W) {
  __shared__ double buf[4096];
  const int p = blockIdx.x;
  if (!W.status[p]) return;
  if (threadIdx.x == 0) atomicAdd(W.fallbacks, 1);
  double x = A.x0[p];
  if (threadIdx.x == 0) A.out[p][A.i0] = x;
  for (int base = 0; base < A.nsteps; base += 4096) {
    const int n = min(4096, A.nsteps - base);
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

int fold_exact(const DenseParams &P, const DenseBuffers &B, const FoldWork &W, int nblk_cap, int dir, int nsteps,
               hipStream_t s) {
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
  hipLaunchKernelGGL(k_fold_sums, dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_plan, dim3(3), dim3(1024), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_local<false>, dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_chain, dim3(3), dim3(kChainThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_local<true>, dim3(A.nblk, 3), dim3(kFoldThreads), 0, s, A, W);
  hipLaunchKernelGGL(k_fold_serial, dim3(3), dim3(256), 0, s, A, W);
  return 0;
}

// per-waveform counters (non-positive Phidot samples, fold fallbacks); a
// kernel rather than a memset so the sequence can be captured in a graph
__global__ void k_clear_flags(int *flags) {
  if (threadIdx.x < 4) flags[threadIdx.x] = 0;
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

__device__ inline double nrm2_scaled(const double *v) {
  double scale = 0.0, ssq = 1.0;
  for (int i = 0; i < 3; i++) {
    const double x = v[i];
    if (x != 0.0) {
      const double ax = fabs(x);
      if (scale < ax) {
        ssq = 1.0 + ssq * (scale / ax) * (scale / ax);
        scale = ax;
      } else {
        ssq += (ax / scale) * (ax / scale);
      }
    }
  }
  return scale * sqrt(ssq);
}

// Per-sample part of the mode sum (AAK.cc:425-481): everything that does not
// depend on the harmonic, folded into the weights of a_n, b_n, c_n in h_I and
// h_II, so that h = Amp * sum_n [cos(nPhi)(alpha_n W_a + gamma_n W_c) +
// sin(nPhi) beta_n W_b] with a_n = Amp alpha_n cos(nPhi) etc. (AAK.cc:493-504).
struct SampleGeom {
  double e, Phi, nu, gimdot, Amp, ti, cosorbphs;
  double W[6];  // a, b, c weights of h_I, then of h_II
};

__device__ inline void sample_geometry(const DenseParams &P, const DenseBuffers &B, int i, SampleGeom &G) {
  const double ti = sample_time(P, i);
  const double e = B.e[i], v = B.v[i], M = B.M[i], S = B.S[i];
  const double gim = B.phase[0][i], alp = B.phase[2][i];
  G.ti = ti;
  G.e = e;
  G.Phi = B.phase[1][i];
  G.gimdot = B.rate[0][i];
  if (i == 0) {
    G.nu = PN::drdm(v, e, P.coslam, S) / PN::dtdm(v, e, P.coslam, S) / (2. * kPi * M);
  } else {
    G.nu = B.rate[1][i] / 2. / kPi;
  }
  const double coslam = P.coslam, sinlam = P.sinlam;
  const double cosqS = P.cosqS, sinqS = P.sinqS, cosqK = P.cosqK, sinqK = P.sinqK;
  const double cosphiK = P.cosphiK, sinphiK = P.sinphiK, halfsqrt3 = P.halfsqrt3;
  const double cosalp = cos(alp), sinalp = sin(alp);
  const double cosqL = cosqK * coslam + sinqK * sinlam * cosalp;
  const double sinqL = sqrt(1. - cosqL * cosqL);
  const double phiLup = sinqK * sinphiK * coslam - cosphiK * sinlam * sinalp - cosqK * sinphiK * sinlam * cosalp;
  const double phiLdown = sinqK * cosphiK * coslam + sinphiK * sinlam * sinalp - cosqK * cosphiK * sinlam * cosalp;
  const double phiL = atan2(phiLup, phiLdown);
  const double Ldotn = cosqL * cosqS + sinqL * sinqS * cos(phiL - P.phiS);
  const double Ldotn2 = Ldotn * Ldotn;
  const double Sdotn = P.Sdotn;
  const double betaup = -Sdotn + coslam * Ldotn;
  const double betadown = sinqS * P.sin_dphiKS * sinlam * cosalp + (cosqK * Sdotn - cosqS) / sinqK * sinlam * sinalp;
  const double beta = atan2(betaup, betadown);
  const double gam = 2. * (gim + beta);
  const double cos2gam = cos(gam), sin2gam = sin(gam);
  double cosorbphs = 0.0, FplusI, FcrosI, FplusII, FcrosII, h3;
  if (P.mich) {
    const double orbphs = 2. * kPi * ti / kYear;
    cosorbphs = cos(orbphs - P.phiS);
    const double sinorbphs = sin(orbphs - P.phiS);
    const double cosq = .5 * cosqS - halfsqrt3 * sinqS * cosorbphs;
    const double phiw = orbphs + atan2(halfsqrt3 * cosqS + .5 * sinqS * cosorbphs, sinqS * sinorbphs);
    const double psiup = .5 * cosqK - halfsqrt3 * sinqK * cos(orbphs - P.phiK) - cosq * P.psi_c0;
    const double psidown = P.psi_c1 - halfsqrt3 * cos(orbphs) * P.psi_c2 - halfsqrt3 * sin(orbphs) * P.psi_c3;
    const double psi = atan2(psiup, psidown);
    const double cosq1 = .5 * (1. + cosq * cosq);
    const double cos2phi = cos(2. * phiw), sin2phi = sin(2. * phiw);
    const double cos2psi = cos(2. * psi), sin2psi = sin(2. * psi);
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
  G.Amp = pow(PN::omega_phi(v, e, coslam, S, M) * P.Mphys * kMsunSec, 2. / 3.) * P.zeta;
  // rotation to the wave frame (KSParMap.cc:532-574); n-independent
  double L[3], nxL[3], nxS[3];
  const double ca = cosalp, sa = sinalp;
  L[0] = P.cosiota_L * P.sin_thK * P.cos_phK + P.siniota_L * (sa * P.sin_phK - ca * P.cos_thK * P.cos_phK);
  L[1] = P.cosiota_L * P.sin_thK * P.sin_phK - P.siniota_L * (sa * P.cos_phK + ca * P.cos_thK * P.sin_phK);
  L[2] = P.cosiota_L * P.cos_thK + P.siniota_L * ca * P.sin_thK;
  const double *n = P.n_vec, *Sv = P.S_vec;
  nxL[0] = n[1] * L[2] - n[2] * L[1];
  nxL[1] = n[2] * L[0] - n[0] * L[2];
  nxL[2] = n[0] * L[1] - n[1] * L[0];
  nxS[0] = n[1] * Sv[2] - n[2] * Sv[1];
  nxS[1] = n[2] * Sv[0] - n[0] * Sv[2];
  nxS[2] = n[0] * Sv[1] - n[1] * Sv[0];
  const double norm = nrm2_scaled(nxL) * nrm2_scaled(nxS);
  const double cosrot = (0.0 + nxL[0] * nxS[0] + nxL[1] * nxS[1] + nxL[2] * nxS[2]) / norm;
  double sinrot = 0.0 + L[0] * nxS[0] + L[1] * nxS[1] + L[2] * nxS[2];
  sinrot -= 0.0 + Sv[0] * nxL[0] + Sv[1] * nxL[1] + Sv[2] * nxL[2];
  sinrot /= norm;
  const double rot0 = 2. * cosrot * cosrot - 1., rot1 = cosrot * sinrot;
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

// One sample, Bessel functions per harmonic (series or Miller) and one
// sincos per harmonic: used where the Taylor form is not applicable.
__device__ inline void modesum_direct(const DenseParams &P, const DenseBuffers &B, const SampleGeom &G, double *hI,
                                      double *hII) {
  const double e = G.e, sq1me2 = sqrt(1 - e * e), gdpi = G.gimdot / kPi;
  double accI = 0.0, accII = 0.0, J[5];
  for (int nn = 1; nn <= P.nmodes; nn++) {
    const double nPhi = harmonic_phase(P, G, (double)nn, gdpi);
    if (P.bessel_series)
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

// ------------------------------------------------------------------ Taylor form
// Inside one block of samples e varies by a few 1e-5: the amplitude factors
// alpha_n, beta_n, gamma_n are expanded in delta = e - e_ref once per block
// (taylor_amplitudes) and evaluated per sample by Horner.  The order p is the
// smallest with remainder bound 7 n (n delta)^{p+1}/(p+1)! <= 1e-17.
constexpr int kMsThreads = 256;
constexpr int kTaylorModes = 128;

// MIXED: the harmonic loop (rotation recurrence, amplitude polynomials,
// response accumulation) runs in FP32 from FP64-reduced phase factors and
// FP64 block coefficients; phases, their reduction mod 2 pi, the block
// expansions and the final scaling stay FP64.
// Block setup of the Taylor form, one workgroup per mode-sum block: the range
// of e over the block's signal samples, the order p and the expansion of
// alpha_n, beta_n, gamma_n for every harmonic (one thread each), written to
// B.ms_coef[block][n][3][kTaylorMax+1]; B.ms_p[block] = p (-1: per-sample path).
constexpr int kSetupThreads = 128;

__global__ __launch_bounds__(kSetupThreads) void k_modesum_setup(DenseParams P, DenseBuffers B) {
  __shared__ double s_sq[kTaylorMax + 1];
  __shared__ double s_min[kSetupThreads / 64], s_max[kSetupThreads / 64];
  __shared__ double s_eref;
  __shared__ int s_p;
  const int tid = threadIdx.x, blk = blockIdx.x;
  const int n_end = min(P.N, P.ip + P.ib);
  double emin = 1e300, emax = -1e300;
  for (int q = tid; q < kMsThreads; q += kSetupThreads) {
    const int i = blk * kMsThreads + q;
    if (i < n_end) {
      const double e = B.e[i];
      emin = fmin(emin, e);
      emax = fmax(emax, e);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    emin = fmin(emin, __shfl_xor(emin, o));
    emax = fmax(emax, __shfl_xor(emax, o));
  }
  if ((tid & 63) == 0) {
    s_min[tid >> 6] = emin;
    s_max[tid >> 6] = emax;
  }
  __syncthreads();
  if (tid == 0) {
    double lo = s_min[0], hi = s_max[0];
    for (int w = 1; w < kSetupThreads / 64; w++) {
      lo = fmin(lo, s_min[w]);
      hi = fmax(hi, s_max[w]);
    }
    int p = -1;
    double eref = 0.0;
    if (lo <= hi && P.bessel_series && P.nmodes <= kTaylorModes && hi <= kSeriesMaxE) {
      eref = 0.5 * (lo + hi);
      p = taylor_order(P.nmodes, 0.5 * (hi - lo));
      if (p >= 0) {
        double sq[kTaylorMax + 1];
        taylor_sqrt1me2(eref, p, sq);
        for (int j = 0; j <= kTaylorMax; j++) s_sq[j] = j <= p ? sq[j] : 0.0;
      }
    }
    s_eref = eref;
    s_p = p;
    B.ms_p[blk] = p;
    B.ms_eref[blk] = eref;
  }
  __syncthreads();
  const int p = s_p;
  if (p >= 0 && tid < P.nmodes) {
    // series length: the host bound for orders n+1, n+2 covers the higher orders
    double coef[3 * (kTaylorMax + 1)];
    taylor_amplitudes(tid + 1, p, s_eref, s_sq, B.bessel_c, B.bessel_lgf, P.bterms[tid + 1], coef);
    double *dst = B.ms_coef + ((size_t)blk * P.nmodes + tid) * (3 * (kTaylorMax + 1));
    for (int q = 0; q < 3 * (kTaylorMax + 1); q++) dst[q] = coef[q];
  }
}

int modesum_blocks(const DenseParams &P) { return (P.N + 2 + kMsThreads - 1) / kMsThreads; }

// MIXED: the harmonic loop (rotation recurrence, amplitude polynomials,
// response accumulation) runs in FP32 from FP64-reduced phase factors and
// FP64 block coefficients; phases, their reduction mod 2 pi, the block
// expansions and the final scaling stay FP64.
template <bool MIXED>
__global__ __launch_bounds__(kMsThreads) void k_modesum(DenseParams P, DenseBuffers B) {
  __shared__ double s_coef[MIXED ? 1 : kTaylorModes][3][kTaylorMax + 1];
  __shared__ float s_coef_f[MIXED ? kTaylorModes : 1][3][kTaylorMax + 1];
  const int tid = threadIdx.x;
  const int i = blockIdx.x * kMsThreads + tid;
  double *hI = B.h, *hII = B.h + (P.N + 2);
  const int n_end = min(P.N, P.ip + P.ib);  // samples [0, n_end) carry the signal
  const bool active = i < n_end;
  const int p = B.ms_p[blockIdx.x];
  const double eref = B.ms_eref[blockIdx.x];
  if (p >= 0) {
    const int nc = P.nmodes * 3 * (kTaylorMax + 1);
    const double *src = B.ms_coef + (size_t)blockIdx.x * nc;
    for (int q = tid; q < nc; q += kMsThreads) {
      if (MIXED) (&s_coef_f[0][0][0])[q] = (float)src[q];
      else (&s_coef[0][0][0])[q] = src[q];
    }
  }
  __syncthreads();
  if (i >= P.N + 2) return;
  if (i >= P.N) {
    B.t[i] = 0.0;
    hI[i] = 0.0;
    hII[i] = 0.0;
    return;
  }
  if (!active) {
    B.t[i] = sample_time(P, i);
    hI[i] = 0.0;
    hII[i] = 0.0;
    return;
  }
  SampleGeom G;
  sample_geometry(P, B, i, G);
  B.t[i] = G.ti;
  double accI, accII;
  if (p < 0) {
    modesum_direct(P, B, G, &accI, &accII);
  } else {
    const double d = G.e - eref;
    // exp(i n Phi) by rotation from Phi reduced mod 2 pi (Cody-Waite, three
    // parts of 2 pi); the Doppler term is linear in n: Doppler_n = n dA + dB
    const double kq = rint(G.Phi * (1.0 / (2.0 * kPi)));
    double phr = fma(-kq, 6.283185307179586, G.Phi);
    phr = fma(-kq, 2.4492935982947064e-16, phr);
    phr = fma(-kq, -5.989539619436679e-33, phr);
    const double gdpi = G.gimdot / kPi;
    double dA = 0.0, dB = 0.0;
    if (P.mich) {
      const double proj = kAUsec * P.sinqS * G.cosorbphs;
      dA = 2. * kPi * G.nu * proj;
      dB = 2. * G.gimdot * proj;
    }
    double ur, ui, qr, qi;
    sincos(phr + dA, &ui, &ur);
    sincos(dB, &qi, &qr);
    accI = 0.0;
    accII = 0.0;
    if (MIXED) {
      float fur = (float)ur, fui = (float)ui, fqr = (float)qr, fqi = (float)qi;
      const float fd = (float)d;
      const float W0 = (float)G.W[0], W1 = (float)G.W[1], W2 = (float)G.W[2];
      const float W3 = (float)G.W[3], W4 = (float)G.W[4], W5 = (float)G.W[5];
      float fI = 0.f, fII = 0.f;
      for (int nn = 1; nn <= P.nmodes; nn++) {
        const float tr = fqr * fur - fqi * fui;
        fqi = fqr * fui + fqi * fur;
        fqr = tr;
        const float *cf = &s_coef_f[nn - 1][0][0];
        float al = cf[p], be = cf[kTaylorMax + 1 + p], ga = cf[2 * (kTaylorMax + 1) + p];
        for (int j = p - 1; j >= 0; j--) {
          al = fmaf(al, fd, cf[j]);
          be = fmaf(be, fd, cf[kTaylorMax + 1 + j]);
          ga = fmaf(ga, fd, cf[2 * (kTaylorMax + 1) + j]);
        }
        fI += fqr * (al * W0 + ga * W2) + fqi * (be * W1);
        fII += fqr * (al * W3 + ga * W5) + fqi * (be * W4);
      }
      accI = (double)fI;
      accII = (double)fII;
    } else {
    for (int nn = 1; nn <= P.nmodes; nn++) {
      // q = exp(i (n (Phi + dA) + dB))
      const double tr = qr * ur - qi * ui;
      qi = qr * ui + qi * ur;
      qr = tr;
      // the reference's rounded argument s = rn(rn(n Phi) + Doppler) differs
      // from n Phi + Doppler by -(e1 + e2), both exact: e1 = n Phi - rn(n Phi),
      // e2 = (rn(n Phi) + Doppler) - s
      const double dn = (double)nn;
      const double pn = dn * G.Phi;
      const double e1 = fma(dn, G.Phi, -pn);
      double s_ = pn, e2 = 0.0;
      if (P.mich) {
        const double fn = dn * G.nu + gdpi;
        const double Doppler = 2. * kPi * fn * kAUsec * P.sinqS * G.cosorbphs;
        s_ = pn + Doppler;
        const double bb = s_ - pn;
        e2 = (pn - (s_ - bb)) + (Doppler - bb);
      }
      const double del = -(e1 + e2);
      const double cnp = qr - del * qi, snp = qi + del * qr;
      double abg[3];
      taylor_eval(&s_coef[nn - 1][0][0], p, d, abg);
      const double al = abg[0], be = abg[1], ga = abg[2];
      accI += cnp * (al * G.W[0] + ga * G.W[2]) + snp * (be * G.W[1]);
      accII += cnp * (al * G.W[3] + ga * G.W[5]) + snp * (be * G.W[4]);
    }
    }
  }
  const double den = taper(P, i, G.ti);
  hI[i] = G.Amp * accI / den;
  hII[i] = G.Amp * accII / den;
}

void launch_modesum(const DenseParams &P, const DenseBuffers &B, hipStream_t s) {
  const int n = P.N + 2;
  hipLaunchKernelGGL(k_modesum_setup, dim3(modesum_blocks(P)), dim3(kSetupThreads), 0, s, P, B);
  if (P.mixed)
    hipLaunchKernelGGL(k_modesum<true>, dim3((n + kMsThreads - 1) / kMsThreads), dim3(kMsThreads), 0, s, P, B);
  else
    hipLaunchKernelGGL(k_modesum<false>, dim3((n + kMsThreads - 1) / kMsThreads), dim3(kMsThreads), 0, s, P, B);
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

__global__ __launch_bounds__(kDotThreads) void k_dot_final(const double *partials, int nparts, double *out) {
  __shared__ double s_dh[kDotThreads], s_hh[kDotThreads];
  double dh = 0.0, hh = 0.0;
  for (int q = threadIdx.x; q < nparts; q += blockDim.x) {
    dh += partials[2 * q];
    hh += partials[2 * q + 1];
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
    out[0] = 4.0 * s_dh[0];
    out[1] = 4.0 * s_hh[0];
  }
}

void launch_weighted_dot(const double2 *hf, const double2 *data, const double *w, int nbins, double2 *hw_out,
                         double *partials, double *out, hipStream_t s) {
  const int need = (2 * nbins + kDotThreads - 1) / kDotThreads;
  const int nb = need < kDotBlocks ? need : kDotBlocks;
  hipLaunchKernelGGL(k_weighted_dot, dim3(nb), dim3(kDotThreads), 0, s, hf, data, w, nbins, hw_out, partials);
  hipLaunchKernelGGL(k_dot_final, dim3(1), dim3(kDotThreads), 0, s, partials, nb, out);
}

}  // namespace aak
