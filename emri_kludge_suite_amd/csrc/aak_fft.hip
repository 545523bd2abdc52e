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
// Each FFT runs in LDS as a Stockham autosort transform (radix 4, one radix-2
// stage when log2 of the length is odd); twiddles come from tables computed on
// the host in extended precision.
#include <hip/hip_runtime.h>

#include "aak_fft.hpp"

namespace aak {

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// One Stockham stage of radix RS on L points held in x (LDS), NT threads.
template <int L, int NT, int NS, int RS>
__device__ inline void fft_stage(double2 *x, int t, const double2 *tw) {
  constexpr int NB = L / RS;                  // butterflies
  constexpr int PER = (NB + NT - 1) / NT;     // per thread
  double2 v[PER][RS];
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = t + u * NT;
    if (NB % NT == 0 || j < NB) {
      const int jm = j & (NS - 1);
#pragma unroll
      for (int r = 0; r < RS; r++) {
        double2 a = x[j + r * NB];
        if (NS > 1 && r > 0) a = cmul(a, tw[(r * jm * (L / (RS * NS))) & (L - 1)]);
        v[u][r] = a;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < PER; u++) {
    const int j = t + u * NT;
    if (NB % NT == 0 || j < NB) {
      const int jm = j & (NS - 1);
      const int d = (j / NS) * NS * RS + jm;
      if (RS == 4) {
        const double2 a0 = make_double2(v[u][0].x + v[u][2].x, v[u][0].y + v[u][2].y);
        const double2 a1 = make_double2(v[u][0].x - v[u][2].x, v[u][0].y - v[u][2].y);
        const double2 a2 = make_double2(v[u][1].x + v[u][3].x, v[u][1].y + v[u][3].y);
        // (v1 - v3) * (-i)
        const double2 a3 = make_double2(v[u][1].y - v[u][3].y, v[u][3].x - v[u][1].x);
        x[d] = make_double2(a0.x + a2.x, a0.y + a2.y);
        x[d + NS] = make_double2(a1.x + a3.x, a1.y + a3.y);
        x[d + 2 * NS] = make_double2(a0.x - a2.x, a0.y - a2.y);
        x[d + 3 * NS] = make_double2(a1.x - a3.x, a1.y - a3.y);
      } else {
        x[d] = make_double2(v[u][0].x + v[u][1].x, v[u][0].y + v[u][1].y);
        x[d + NS] = make_double2(v[u][0].x - v[u][1].x, v[u][0].y - v[u][1].y);
      }
    }
  }
  __syncthreads();
}

template <int L, int NT, int NS>
__device__ inline void fft_stages(double2 *x, int t, const double2 *tw) {
  if constexpr (NS < L) {
    constexpr int RS = ((L / NS) % 4 == 0) ? 4 : 2;
    fft_stage<L, NT, NS, RS>(x, t, tw);
    fft_stages<L, NT, NS * RS>(x, t, tw);
  }
}

// Forward L-point DFT of x in place (natural order in and out); tw = W_L^e.
// Every thread of the block must call it (it synchronises the block).
template <int L, int NT>
__device__ inline void fft_lds(double2 *x, int t, const double2 *tw) {
  fft_stages<L, NT, 1>(x, t, tw);
}

__device__ inline double2 tw2(const double2 *hi, const double2 *lo, long long e) {
  return cmul(hi[e >> 10], lo[e & 1023]);
}

// ---------------------------------------------------------------- pass 1
// columns per workgroup: 8 (one 128-byte line per row) while 8 columns fit LDS
template <int LR>
constexpr int col_cols() { return LR <= 1024 ? 8 : 4; }
constexpr int kColNT = 64;      // threads per column

template <int LR>
__global__ __launch_bounds__(col_cols<LR>() * kColNT) void k_fft_cols(FftArgs F, const double *h, double2 *Y) {
  constexpr int kColCols = col_cols<LR>();
  constexpr int LP = LR + 2;
  __shared__ double2 s[kColCols * LP];
  const int ch = blockIdx.y, b0 = blockIdx.x * kColCols, tid = threadIdx.x;
  const double2 *z = reinterpret_cast<const double2 *>(h + (size_t)ch * (F.N + 2));
  for (int idx = tid; idx < kColCols * LR; idx += kColCols * kColNT) {
    const int a = idx / kColCols, c = idx % kColCols;
    s[c * LP + a] = z[(size_t)a * F.C + b0 + c];
  }
  __syncthreads();
  fft_lds<LR, kColNT>(s + (tid / kColNT) * LP, tid % kColNT, F.twR);
  double2 *Yc = Y + (size_t)ch * F.M;
  for (int idx = tid; idx < kColCols * LR; idx += kColCols * kColNT) {
    const int k1 = idx / kColCols, c = idx % kColCols, b = b0 + c;
    Yc[(size_t)k1 * F.C + b] = cmul(s[c * LP + k1], tw2(F.twMhi, F.twMlo, (long long)b * k1));
  }
}

// ---------------------------------------------------------------- pass 2
constexpr int kRowThreads = 256;

template <int LC>
__global__ __launch_bounds__(kRowThreads) void k_fft_rows(FftArgs F, const double2 *Y, double *partials,
                                                          double2 *hw_out) {
  constexpr int LP = LC + 2;
  constexpr int NT = kRowThreads / 2;
  __shared__ double2 s[2 * LP];
  __shared__ double s_red[2][kRowThreads / 64];
  const int ch = blockIdx.y, p = blockIdx.x, tid = threadIdx.x;
  const int A = p, B = (F.R - p) & (F.R - 1);
  const double2 *Yc = Y + (size_t)ch * F.M;
  for (int idx = tid; idx < 2 * LC; idx += kRowThreads) {
    const int half = idx / LC, b = idx % LC;
    s[half * LP + b] = Yc[(size_t)(half ? B : A) * F.C + b];
  }
  __syncthreads();
  fft_lds<LC, NT>(s + (tid / NT) * LP, tid % NT, F.twC);
  const double2 *sA = s, *sB = s + LP;
  const double2 *Dp = F.Dp[ch];
  const double *wp = F.wp[ch];
  double2 *hw = hw_out ? hw_out + (size_t)ch * (F.M + 1) : nullptr;
  double dh = 0.0, hh = 0.0;
  // H_k from Z_k and conj(Z_(M-k)); pidx: position of k in the stored data order
  auto emit = [&](long long k, double2 zk, double2 zm, int pidx) {
    const double2 zmc = make_double2(zm.x, -zm.y);
    const double2 E = make_double2(0.5 * (zk.x + zmc.x), 0.5 * (zk.y + zmc.y));
    const double2 dz = make_double2(zk.x - zmc.x, zk.y - zmc.y);
    const double2 O = make_double2(0.5 * dz.y, -0.5 * dz.x);
    const double2 WO = cmul(tw2(F.twNhi, F.twNlo, k), O);
    const double w = wp[pidx];
    const double2 x = make_double2((E.x + WO.x) * w, (E.y + WO.y) * w);
    const double2 d = Dp[pidx];
    dh += x.x * d.x + x.y * d.y;
    hh += x.x * x.x + x.y * x.y;
    if (hw) hw[k] = x;
  };
  for (int k2 = tid; k2 < LC; k2 += kRowThreads) {
    if (A == 0) {
      if (k2 == 0) {
        // H_0 = E_0 + O_0 and H_M = E_0 - O_0, both real
        const double2 z0 = sA[0];
        const double w0 = wp[0], wm = wp[F.M];
        const double2 x0 = make_double2((z0.x + z0.y) * w0, 0.0), xm = make_double2((z0.x - z0.y) * wm, 0.0);
        dh += x0.x * Dp[0].x + x0.y * Dp[0].y + xm.x * Dp[F.M].x + xm.y * Dp[F.M].y;
        hh += x0.x * x0.x + xm.x * xm.x;
        if (hw) {
          hw[0] = x0;
          hw[F.M] = xm;
        }
      } else {
        emit((long long)F.R * k2, sA[k2], sA[LC - k2], k2);
      }
    } else if (A == B) {
      emit(A + (long long)F.R * k2, sA[k2], sA[LC - 1 - k2], A * LC + k2);
    } else {
      emit(A + (long long)F.R * k2, sA[k2], sB[LC - 1 - k2], A * LC + k2);
      emit(B + (long long)F.R * k2, sB[k2], sA[LC - 1 - k2], B * LC + k2);
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    dh += __shfl_xor(dh, o, 64);
    hh += __shfl_xor(hh, o, 64);
  }
  if ((tid & 63) == 0) {
    s_red[0][tid >> 6] = dh;
    s_red[1][tid >> 6] = hh;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, b = 0.0;
    for (int w = 0; w < kRowThreads / 64; w++) {
      a += s_red[0][w];
      b += s_red[1][w];
    }
    const int q = ch * (F.R / 2 + 1) + p;
    partials[2 * q] = a;
    partials[2 * q + 1] = b;
  }
}

// ---------------------------------------------------------------- launch
template <int LR>
static void launch_cols(const FftArgs &F, const double *h, double2 *Y, hipStream_t s) {
  hipLaunchKernelGGL(k_fft_cols<LR>, dim3(F.C / col_cols<LR>(), 2), dim3(col_cols<LR>() * kColNT), 0, s, F, h, Y);
}

template <int LC>
static void launch_rows(const FftArgs &F, const double2 *Y, double *partials, double2 *hw, hipStream_t s) {
  hipLaunchKernelGGL(k_fft_rows<LC>, dim3(F.R / 2 + 1, 2), dim3(kRowThreads), 0, s, F, Y, partials, hw);
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

int fft_partials(const FftArgs &F) { return 2 * (F.R / 2 + 1); }

void launch_fft_pass1(const FftArgs &F, const double *h, double2 *Y, hipStream_t s) {
  switch (F.R) {
    case 16: launch_cols<16>(F, h, Y, s); break;
    case 32: launch_cols<32>(F, h, Y, s); break;
    case 64: launch_cols<64>(F, h, Y, s); break;
    case 128: launch_cols<128>(F, h, Y, s); break;
    case 256: launch_cols<256>(F, h, Y, s); break;
    case 512: launch_cols<512>(F, h, Y, s); break;
    case 1024: launch_cols<1024>(F, h, Y, s); break;
    default: launch_cols<2048>(F, h, Y, s); break;
  }
}

void launch_fft_pass2(const FftArgs &F, const double2 *Y, double *partials, double2 *hw_out, hipStream_t s) {
  switch (F.C) {
    case 8: launch_rows<8>(F, Y, partials, hw_out, s); break;
    case 16: launch_rows<16>(F, Y, partials, hw_out, s); break;
    case 32: launch_rows<32>(F, Y, partials, hw_out, s); break;
    case 64: launch_rows<64>(F, Y, partials, hw_out, s); break;
    case 128: launch_rows<128>(F, Y, partials, hw_out, s); break;
    case 256: launch_rows<256>(F, Y, partials, hw_out, s); break;
    case 512: launch_rows<512>(F, Y, partials, hw_out, s); break;
    case 1024: launch_rows<1024>(F, Y, partials, hw_out, s); break;
    default: launch_rows<2048>(F, Y, partials, hw_out, s); break;
  }
}

}  // namespace aak
