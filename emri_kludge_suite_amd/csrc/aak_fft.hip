// aak_fft.hip -- noise-weighted inner products of a real waveform pair on
// CDNA4 (gfx950): a real FFT of length N = 2M fused with the PSD weighting and
// the two reductions <h w, d w> and <h w, h w> (AAK_manager.cu:352-446 and
// kernel.cu:557-567 semantics: one-sided spectrum, DC and Nyquist bins with
// full weight).
//
// The M-point complex FFT of z_m = h_2m + i h_2m+1 runs in four-step form with
// M = R C, m = a C + b, k = k1 + R k2:
//   pass 1 (k_fft_cols):  R-point FFTs down the columns b, times W_M^(b k1);
//   pass 2 (k_fft_rows):  C-point FFTs along the rows k1.
// Row k1 and row R - k1 hold Z_k and Z_(M-k) for every k of either row, so the
// second pass also forms the real spectrum H_k = E_k + W_N^k O_k,
// E_k = (Z_k + conj Z_(M-k)) / 2, O_k = (Z_k - conj Z_(M-k)) / 2i, weights it
// and accumulates the two sums; the spectrum is written only on request.  The
// weighted data and weights are stored once, at input time, in that pass's
// order (row k1, then k2), so every read is contiguous.  HBM traffic per
// likelihood is about 2 x (16 + 16 + 16 + 16 + 8) MB at N = 2^21.
//
// Each FFT is a Stockham autosort transform of radix-16 stages (a radix-8,
// -4 or -2 stage finishes lengths that are not powers of 16) on registers,
// exchanging through LDS between stages; LDS rows carry one pad element per 16
// so that the stride-16 writes of the first stage spread over all banks.
// Twiddles come from tables computed on the host in extended precision.
#include <hip/hip_runtime.h>

#include "aak_fft.hpp"

namespace aak {

__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }  // a * (-i)

// cos(pi/8), sin(pi/8), sqrt(1/2)
constexpr double kC8 = 0.92387953251128675613;
constexpr double kS8 = 0.38268343236508977173;
constexpr double kR2 = 0.70710678118654752440;

__device__ inline void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3) {
  const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2), t2 = cadd(a1, a3), t3 = cmul_mi(csub(a1, a3));
  a0 = cadd(t0, t2);
  a1 = cadd(t1, t3);
  a2 = csub(t0, t2);
  a3 = csub(t1, t3);
}

// Forward DFT of RS points, v in natural order, o out in natural order.
template <int RS>
__device__ inline void dft(double2 *v, double2 *o) {
  if constexpr (RS == 2) {
    o[0] = cadd(v[0], v[1]);
    o[1] = csub(v[0], v[1]);
  } else if constexpr (RS == 4) {
    dft4(v[0], v[1], v[2], v[3]);
#pragma unroll
    for (int r = 0; r < 4; r++) o[r] = v[r];
  } else if constexpr (RS == 8) {
    // n = 2 n1 + n2, k = k1 + 4 k2
    dft4(v[0], v[2], v[4], v[6]);
    dft4(v[1], v[3], v[5], v[7]);
    v[3] = cmul(v[3], make_double2(kR2, -kR2));
    v[5] = cmul_mi(v[5]);
    v[7] = cmul(v[7], make_double2(-kR2, -kR2));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) {
      o[k1] = cadd(v[2 * k1], v[2 * k1 + 1]);
      o[k1 + 4] = csub(v[2 * k1], v[2 * k1 + 1]);
    }
  } else {
    static_assert(RS == 16, "radix 2, 4, 8 or 16");
    // n = 4 n1 + n2, k = k1 + 4 k2
#pragma unroll
    for (int n2 = 0; n2 < 4; n2++) dft4(v[n2], v[4 + n2], v[8 + n2], v[12 + n2]);
    // v[n2 + 4 k1] *= W16^(n2 k1)
    v[5] = cmul(v[5], make_double2(kC8, -kS8));
    v[9] = cmul(v[9], make_double2(kR2, -kR2));
    v[13] = cmul(v[13], make_double2(kS8, -kC8));
    v[6] = cmul(v[6], make_double2(kR2, -kR2));
    v[10] = cmul_mi(v[10]);
    v[14] = cmul(v[14], make_double2(-kR2, -kR2));
    v[7] = cmul(v[7], make_double2(kS8, -kC8));
    v[11] = cmul(v[11], make_double2(-kR2, -kR2));
    v[15] = cmul(v[15], make_double2(-kC8, kS8));
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++) dft4(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
#pragma unroll
    for (int k1 = 0; k1 < 4; k1++)
#pragma unroll
      for (int k2 = 0; k2 < 4; k2++) o[k1 + 4 * k2] = v[4 * k1 + k2];
  }
}

constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }
// radix of the stage that faces `rem` remaining points: 16 while that leaves a
// power of 16 or at least 4 behind, then 8, 4 or 2
constexpr int stage_radix(int rem) {
  return (ilog2(rem) >= 4 && ilog2(rem) != 5) ? 16 : (ilog2(rem) == 5 ? 8 : rem);
}

__device__ inline int pad(int i) { return i + (i >> 4); }

// One Stockham stage of radix RS on L points held in x (LDS, padded), NT threads.
template <int L, int NT, int NS, int RS>
__device__ inline void fft_stage(double2 *x, int t, const double2 *tw) {
  constexpr int NB = L / RS;
  constexpr int PER = (NB + NT - 1) / NT;
  double2 o[PER][RS];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = t + u * NT;
    if (NB % NT == 0 || j < NB) {
      const int jm = j & (NS - 1);
      double2 v[RS];
#pragma unroll
      for (int r = 0; r < RS; r++) v[r] = x[pad(j + r * NB)];
      if constexpr (NS > 1) {
#pragma unroll
        for (int r = 1; r < RS; r++) v[r] = cmul(v[r], tw[(r * jm * (L / (RS * NS))) & (L - 1)]);
      }
      dft<RS>(v, o[u]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = t + u * NT;
    if (NB % NT == 0 || j < NB) {
      const int d = (j / NS) * NS * RS + (j & (NS - 1));
#pragma unroll
      for (int r = 0; r < RS; r++) x[pad(d + r * NS)] = o[u][r];
    }
  }
  __syncthreads();
}

template <int L, int NT, int NS>
__device__ inline void fft_stages(double2 *x, int t, const double2 *tw) {
  if constexpr (NS < L) {
    constexpr int RS = stage_radix(L / NS);
    fft_stage<L, NT, NS, RS>(x, t, tw);
    fft_stages<L, NT, NS * RS>(x, t, tw);
  }
}

// Whole L-point FFT whose first stage (no twiddles) reads straight from global
// memory: butterfly j reads src[j + r L / RS], contiguous across the threads
// for each r.
template <int L, int NT>
__device__ inline void fft_first_global(const double2 *src, double2 *x, int t, const double2 *tw) {
  constexpr int RS = stage_radix(L);
  constexpr int NB = L / RS;
  constexpr int PER = (NB + NT - 1) / NT;
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = t + u * NT;
    if (NB % NT == 0 || j < NB) {
      double2 v[RS], o[RS];
#pragma unroll
      for (int r = 0; r < RS; r++) v[r] = src[j + r * NB];
      dft<RS>(v, o);
#pragma unroll
      for (int r = 0; r < RS; r++) x[pad(j * RS + r)] = o[r];
    }
  }
  __syncthreads();
  fft_stages<L, NT, RS>(x, t, tw);
}

__device__ inline double2 tw2(const double2 *hi, const double2 *lo, long long e) {
  return cmul(hi[e >> 10], lo[e & 1023]);
}

// ---------------------------------------------------------------- pass 1
// columns per workgroup: 8 (one 128-byte line per row) up to 512-point
// columns, 4 beyond (two workgroups per CU at 1024 points); the workgroups
// of the two halves of a line are placed on one XCD (col_group)
template <int LR>
constexpr int col_cols() { return LR <= 512 ? 8 : 4; }
// LDS stride of a column: padded length, plus one so the columns of a row
// start on different banks
template <int LR>
constexpr int col_stride() { return LR + LR / 16 + 1; }
// threads per column or row: one radix-16 butterfly each in the first stage
template <int L>
constexpr int fft_threads() { return L / 16 > 64 ? L / 16 : 64; }

template <int LR>
__global__ __launch_bounds__(col_cols<LR>() * fft_threads<LR>()) void k_fft_cols(FftArgs F, const double *h0,
                                                                                  const double *h1, double2 *Y) {
  constexpr int CC = col_cols<LR>(), CS = col_stride<LR>(), NT = fft_threads<LR>();
  __shared__ double2 s[CC * CS];
  int ch = blockIdx.y, grp = blockIdx.x;
  if (CC < 8) {
    // workgroups are dealt round-robin to the 8 XCDs in launch order; pair
    // groups 2q and 2q+1 (one 128-byte line per row between them) on one XCD,
    // launched 8 apart, so each line is fetched into one L2 once
    const int L = blockIdx.y * gridDim.x + blockIdx.x, total = gridDim.x * gridDim.y;
    if ((total & 15) == 0) {
      const int xcd = L & 7, r = L >> 3;
      const int g = ((((r >> 1) << 3) | xcd) << 1) | (r & 1);
      ch = g / gridDim.x;
      grp = g % gridDim.x;
    }
  }
  const int b0 = grp * CC, tid = threadIdx.x;
  constexpr int PT = (LR + NT - 1) / NT;  // elements per thread in the staging loops
  constexpr bool kFull = LR % NT == 0;
  const double2 *z = reinterpret_cast<const double2 *>(ch ? h1 : h0);
  {
    // every load in flight before the first LDS store
    double2 v[PT];
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int idx = tid + q * CC * NT;
      if (kFull || idx < CC * LR) v[q] = z[(size_t)(idx / CC) * F.C + b0 + idx % CC];
    }
#pragma unroll
    for (int q = 0; q < PT; q++) {
      const int idx = tid + q * CC * NT;
      if (kFull || idx < CC * LR) s[(idx % CC) * CS + pad(idx / CC)] = v[q];
    }
  }
  __syncthreads();
  fft_stages<LR, NT, 1>(s + (tid / NT) * CS, tid % NT, F.twR);
  double2 *Yc = Y + (size_t)ch * F.M;
#pragma unroll
  for (int q = 0; q < PT; q++) {
    const int idx = tid + q * CC * NT;
    const int k1 = idx / CC, c = idx % CC, b = b0 + c;
    if (kFull || idx < CC * LR)
      Yc[(size_t)k1 * F.C + b] = cmul(s[c * CS + pad(k1)], tw2(F.twMhi, F.twMlo, (long long)b * k1));
  }
}

// ---------------------------------------------------------------- pass 2
// One workgroup per row pair, half of it per row: block 0 takes rows 0 and R/2
// (each its own mirror), block p > 0 rows p and R - p.
constexpr int kFinalLanes = 256;       // lanes of the final reduction's order
constexpr int kFinalSpinCap = 1 << 20;  // polls of one missing pair before it reads as NaN

constexpr int kFinalMaxPer = 8;         // pairs per lane: F.R <= 2048

__device__ inline long long final_peek(const double *p) {
  return __hip_atomic_load(reinterpret_cast<const long long *>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The value of a published word, first read as v (polled again while it still
// reads kFinalEmpty); the word is returned to kFinalEmpty.
__device__ inline double final_take(double *p, long long v) {
  long long *w = reinterpret_cast<long long *>(p);
  int spins = 0;
  while (v == kFinalEmpty) {
    if (++spins > kFinalSpinCap) {
      v = 0x7FF8000000000000LL;
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __hip_atomic_store(w, kFinalEmpty, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double(v);
}

__device__ inline void final_put(double *p, double x) {
  __hip_atomic_store(reinterpret_cast<long long *>(p), __double_as_longlong(x), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_fft_final_reset(double *partials, int nwords) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += gridDim.x * blockDim.x)
    reinterpret_cast<long long *>(partials)[i] = kFinalEmpty;
}

// The final reduction, run by one workgroup of BS threads with LDS scratch
// sd (2 kFinalLanes doubles): collects the R published pairs and sums them in
// the order of k_dot_final.
template <int BS>
__device__ void final_collect(int R, double *partials, const FftFinal &fin, double *sd) {
  const int tid = threadIdx.x;
  double *sh = sd + kFinalLanes;
  // every word of the thread's lanes is requested before the first is used
  constexpr int kVL = (kFinalLanes + BS - 1) / BS;  // lanes per thread
  long long w[kVL][2 * kFinalMaxPer];
#pragma unroll
  for (int v = 0; v < kVL; v++)
#pragma unroll
    for (int k = 0; k < kFinalMaxPer; k++) {
      const int q = tid + v * BS + k * kFinalLanes;
      if (tid + v * BS < kFinalLanes && q < R) {
        w[v][2 * k] = final_peek(partials + 2 * q);
        w[v][2 * k + 1] = final_peek(partials + 2 * q + 1);
      }
    }
#pragma unroll
  for (int v = 0; v < kVL; v++) {
    const int t = tid + v * BS;
    if (t >= kFinalLanes) break;
    double pd = 0.0, ph = 0.0;
#pragma unroll
    for (int k = 0; k < kFinalMaxPer; k++) {
      const int q = t + k * kFinalLanes;
      if (q < R) {
        pd += final_take(partials + 2 * q, w[v][2 * k]);
        ph += final_take(partials + 2 * q + 1, w[v][2 * k + 1]);
      }
    }
    sd[t] = pd;
    sh[t] = ph;
  }
  __syncthreads();
  for (int o = kFinalLanes / 2; o > 0; o >>= 1) {
    for (int t = tid; t < o; t += BS) {
      sd[t] += sd[t + o];
      sh[t] += sh[t + o];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const bool bad = fin.flags && fin.flags[3] != 0;
    fin.out[0] = bad ? __builtin_nan("") : 4.0 * sd[0];
    fin.out[1] = bad ? __builtin_nan("") : 4.0 * sh[0];
  }
  __syncthreads();
  if (fin.clear && tid < 8) fin.clear[tid] = 0;
}

template <int LC, bool FINAL>
__global__ __launch_bounds__(2 * fft_threads<LC>()) void k_fft_rows(FftArgs F, const double2 *Y, double *partials,
                                                                    double2 *hw_out, FftFinal fin) {
  constexpr int RSD = LC + LC / 16, NT = fft_threads<LC>();
  __shared__ double2 s[2 * RSD];
  __shared__ double s_red[2 * NT / 64][2];
  const int ch = blockIdx.y, blk = blockIdx.x, tid = threadIdx.x;
  if (FINAL && blk == F.R / 2) {
    // the extra column of workgroups: (R/2, 1), dispatched after every row
    // workgroup, collects their pairs, so each pair it waits for is on its
    // way; (R/2, 0) has nothing to do.  LDS of the FFT serves as scratch where
    // it is large enough.
    constexpr bool kReuse = 4 * RSD >= 2 * kFinalLanes;
    __shared__ double s_fin[kReuse ? 1 : 2 * kFinalLanes];
    if (ch == 1) final_collect<2 * NT>(F.R, partials, fin, kReuse ? reinterpret_cast<double *>(s) : s_fin);
    return;
  }
  const int half = tid / NT;
  const int A = blk, B = blk == 0 ? F.R / 2 : F.R - blk;
  const double2 *Yc = Y + (size_t)ch * F.M;
  const double2 *Dp = F.Dp[ch];
  const double *wp = F.wp[ch];
  // this thread's weighted data and weights, fetched before the FFT so their
  // latency hides behind it; k2 = tid + q 2NT in rows A and B
  constexpr int KP = (LC + 2 * NT - 1) / (2 * NT);
  constexpr bool kFull = LC % (2 * NT) == 0;
  double2 dA[KP], dB[KP];
  double wA[KP], wB[KP];
#pragma unroll
  for (int q = 0; q < KP; q++) {
    const int k2 = min(tid + q * 2 * NT, LC - 1);
    dA[q] = Dp[A * LC + k2];
    wA[q] = wp[A * LC + k2];
    dB[q] = Dp[B * LC + k2];
    wB[q] = wp[B * LC + k2];
  }
  fft_first_global<LC, NT>(Yc + (size_t)(half ? B : A) * F.C, s + half * RSD, tid % NT, F.twC);
  const double2 *sA = s, *sB = s + RSD;
  double2 *hw = hw_out ? hw_out + (size_t)ch * (F.M + 1) : nullptr;
  double dh = 0.0, hh = 0.0;
  // W_N^(row + R k2) = W_N^row W_N^(R k2)
  const double2 WrA = tw2(F.twNhi, F.twNlo, A), WrB = tw2(F.twNhi, F.twNlo, B);
  // H_k from Z_k and conj(Z_(M-k)), k = row + R k2, weighted and summed
  auto emit = [&](long long k, double2 W, double2 zk, double2 zm, double2 d, double w) {
    const double2 zmc = make_double2(zm.x, -zm.y);
    const double2 E = make_double2(0.5 * (zk.x + zmc.x), 0.5 * (zk.y + zmc.y));
    const double2 dz = make_double2(zk.x - zmc.x, zk.y - zmc.y);
    const double2 O = make_double2(0.5 * dz.y, -0.5 * dz.x);
    const double2 WO = cmul(W, O);
    const double2 x = make_double2((E.x + WO.x) * w, (E.y + WO.y) * w);
    dh += x.x * d.x + x.y * d.y;
    hh += x.x * x.x + x.y * x.y;
    if (hw) hw[k] = x;
  };
#pragma unroll
  for (int q = 0; q < KP; q++) {
    const int k2 = tid + q * 2 * NT;
    if (kFull || k2 < LC) {
      const double2 t = F.twNr[k2];
      const long long kA = A + (long long)F.R * k2, kB = B + (long long)F.R * k2;
      if (blk == 0) {
        // row 0 (Z_(M-k) in row 0 itself) and row R/2 (its own mirror)
        if (k2 == 0) {
          // H_0 = E_0 + O_0 and H_M = E_0 - O_0, both real
          const double2 z0 = sA[0];
          const double wm = wp[F.M];
          const double2 dm = Dp[F.M];
          const double2 x0 = make_double2((z0.x + z0.y) * wA[q], 0.0), xm = make_double2((z0.x - z0.y) * wm, 0.0);
          dh += x0.x * dA[q].x + x0.y * dA[q].y + xm.x * dm.x + xm.y * dm.y;
          hh += x0.x * x0.x + xm.x * xm.x;
          if (hw) {
            hw[0] = x0;
            hw[F.M] = xm;
          }
        } else {
          emit(kA, t, sA[pad(k2)], sA[pad(LC - k2)], dA[q], wA[q]);
        }
        emit(kB, cmul(WrB, t), sB[pad(k2)], sB[pad(LC - 1 - k2)], dB[q], wB[q]);
      } else {
        emit(kA, cmul(WrA, t), sA[pad(k2)], sB[pad(LC - 1 - k2)], dA[q], wA[q]);
        emit(kB, cmul(WrB, t), sB[pad(k2)], sA[pad(LC - 1 - k2)], dB[q], wB[q]);
      }
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    dh += __shfl_xor(dh, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  if ((tid & 63) == 0) {
    s_red[tid >> 6][0] = dh;
    s_red[tid >> 6][1] = hh;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < 2 * NT / 64; w++) {
      a += s_red[w][0];
      b += s_red[w][1];
    }
    const int q = ch * (F.R / 2) + blk;
    if (FINAL) {
      final_put(partials + 2 * q, a);
      final_put(partials + 2 * q + 1, b);
    } else {
      partials[2 * q] = a;
      partials[2 * q + 1] = b;
    }
  }
}

// ---------------------------------------------------------------- launch
template <int LR>
static void launch_cols(const FftArgs &F, const double *h0, const double *h1, double2 *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_fft_cols<LR>, dim3(F.C / col_cols<LR>(), 2), dim3(col_cols<LR>() * fft_threads<LR>()), 0, s, F,
                     h0, h1, Y);
}

template <int LC>
static void launch_rows(const FftArgs &F, const double2 *Y, double *partials, double2 *hw, const FftFinal *fin,
                        hipStream_t s) {
  if (fin)
    hipLaunchKernelGGL((k_fft_rows<LC, true>), dim3(F.R / 2 + 1, 2), dim3(2 * fft_threads<LC>()), 0, s, F, Y,
                       partials, nullptr, *fin);
  else
    hipLaunchKernelGGL((k_fft_rows<LC, false>), dim3(F.R / 2, 2), dim3(2 * fft_threads<LC>()), 0, s, F, Y, partials,
                       hw, FftFinal{});
}

bool fft_supported(int N, int *R, int *C) {
  if (N < 256 || (N & (N - 1))) return false;
  int m = 0;
  while ((1 << m) < N / 2) m++;
  if (m < 7 || m > 22) return false;
  *R = 1 << ((m + 1) / 2);
  *C = (N / 2) / *R;
  return true;
}

int fft_partials(const FftArgs &F) { return F.R; }

// Table layout: W_R | W_C | W_M lo | W_M hi | W_N lo | W_N hi | W_N^(R k2).
static void table_offsets(const FftArgs &F, size_t o[8]) {
  const size_t nhi = (size_t)F.M / 1024 + 2;
  o[0] = 0;
  o[1] = o[0] + F.R;
  o[2] = o[1] + F.C;
  o[3] = o[2] + 1024;
  o[4] = o[3] + nhi;
  o[5] = o[4] + 1024;
  o[6] = o[5] + nhi;
  o[7] = o[6] + F.C;
}

size_t fft_table_size(const FftArgs &F) {
  size_t o[8];
  table_offsets(F, o);
  return o[7];
}

void fft_fill_tables(const FftArgs &F, double2 *tw) {
  // W_L^e = exp(-2 pi i e / L), computed in extended precision
  auto root = [](long long e, long long L) {
    const long double pi = 3.141592653589793238462643383279502884L;
    const long double a = -2.0L * pi * (long double)(e % L) / (long double)L;
    return make_double2((double)cosl(a), (double)sinl(a));
  };
  size_t o[8];
  table_offsets(F, o);
  const int nhi = (int)(o[4] - o[3]);
  for (int e = 0; e < F.R; e++) tw[o[0] + e] = root(e, F.R);
  for (int e = 0; e < F.C; e++) tw[o[1] + e] = root(e, F.C);
  for (int e = 0; e < 1024; e++) {
    tw[o[2] + e] = root(e, F.M);
    tw[o[4] + e] = root(e, F.N);
  }
  for (int e = 0; e < nhi; e++) {
    tw[o[3] + e] = root(1024LL * e, F.M);
    tw[o[5] + e] = root(1024LL * e, F.N);
  }
  for (int e = 0; e < F.C; e++) tw[o[6] + e] = root((long long)F.R * e, F.N);
}

void fft_bind_tables(FftArgs *F, const double2 *tw) {
  size_t o[8];
  table_offsets(*F, o);
  F->twR = tw + o[0];
  F->twC = tw + o[1];
  F->twMlo = tw + o[2];
  F->twMhi = tw + o[3];
  F->twNlo = tw + o[4];
  F->twNhi = tw + o[5];
  F->twNr = tw + o[6];
}

void launch_fft_pass1(const FftArgs &F, const double *h0, const double *h1, double2 *Y, hipStream_t s) {
  switch (F.R) {
    case 16: launch_cols<16>(F, h0, h1, Y, s); break;
    case 32: launch_cols<32>(F, h0, h1, Y, s); break;
    case 64: launch_cols<64>(F, h0, h1, Y, s); break;
    case 128: launch_cols<128>(F, h0, h1, Y, s); break;
    case 256: launch_cols<256>(F, h0, h1, Y, s); break;
    case 512: launch_cols<512>(F, h0, h1, Y, s); break;
    case 1024: launch_cols<1024>(F, h0, h1, Y, s); break;
    default: launch_cols<2048>(F, h0, h1, Y, s); break;
  }
}

static void launch_pass2(const FftArgs &F, const double2 *Y, double *partials, double2 *hw_out, const FftFinal *fin,
                         hipStream_t s) {
  switch (F.C) {
    case 8: launch_rows<8>(F, Y, partials, hw_out, fin, s); break;
    case 16: launch_rows<16>(F, Y, partials, hw_out, fin, s); break;
    case 32: launch_rows<32>(F, Y, partials, hw_out, fin, s); break;
    case 64: launch_rows<64>(F, Y, partials, hw_out, fin, s); break;
    case 128: launch_rows<128>(F, Y, partials, hw_out, fin, s); break;
    case 256: launch_rows<256>(F, Y, partials, hw_out, fin, s); break;
    case 512: launch_rows<512>(F, Y, partials, hw_out, fin, s); break;
    case 1024: launch_rows<1024>(F, Y, partials, hw_out, fin, s); break;
    default: launch_rows<2048>(F, Y, partials, hw_out, fin, s); break;
  }
}

void launch_fft_pass2(const FftArgs &F, const double2 *Y, double *partials, double2 *hw_out, hipStream_t s) {
  launch_pass2(F, Y, partials, hw_out, nullptr, s);
}

void launch_fft_pass2_final(const FftArgs &F, const double2 *Y, double *partials, const FftFinal &fin, hipStream_t s) {
  launch_pass2(F, Y, partials, nullptr, &fin, s);
}

void launch_fft_final_reset(double *partials, int npairs, hipStream_t s) {
  hipLaunchKernelGGL(k_fft_final_reset, dim3(8), dim3(256), 0, s, partials, 2 * npairs);
}

}  // namespace aak
