// aak_noalloc.hip -- the noise-weighted inner product on caller-owned device
// buffers, with no device allocation: aak_inner_product_noalloc.
//
// The kernels are those of the likelihood stage (the fused four-step FFT of
// aak_fft.hip and the final reduction).  Their scratch -- the four-step
// intermediate, the weighted data and weights in the FFT's second-pass order,
// the twiddle tables and the partial sums -- lives in one file-scope device
// array, `scratch`, of kScratchBytes per GPU, held from module load to the
// end of the process.  Its device address is taken once per device with
// hipGetSymbolAddress and used for every kernel argument; the array's name
// itself is never handed to a kernel.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "aak_dense.hpp"
#include "aak_fft.hpp"

namespace aak {
int set_last_error(int code, const std::string &msg);
}

using namespace aak;

namespace {

// Largest length served: N = 2^21 (M = 2^20 complex points, R = C = 1024).
constexpr int kMaxN = 1 << 21;
constexpr size_t kMaxM = (size_t)kMaxN / 2;
constexpr size_t kAlign = 256;

constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

// Regions for the largest length; smaller lengths use prefixes of each.
constexpr size_t kBytesY = align_up(sizeof(double2) * 2 * kMaxM);        // four-step intermediate, 2 channels
constexpr size_t kBytesD = align_up(sizeof(double2) * 2 * (kMaxM + 1));  // weighted data, second-pass order
constexpr size_t kBytesW = align_up(sizeof(double) * 2 * (kMaxM + 1));   // weights, second-pass order
constexpr size_t kBytesTw = align_up(sizeof(double2) * 8192);            // twiddle tables (7172 entries at kMaxN)
constexpr size_t kBytesP = align_up(sizeof(double) * 2 * 2048);          // partial-sum pairs (R <= 2048)
constexpr size_t kOffY = 0, kOffD = kOffY + kBytesY, kOffW = kOffD + kBytesD, kOffTw = kOffW + kBytesW,
                 kOffP = kOffTw + kBytesTw, kScratchBytes = kOffP + kBytesP;  // ~80 MiB

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device
int g_table_n[kMaxDevices] = {};             // length whose twiddles the device's scratch holds

int fail(int code, const std::string &msg) { return set_last_error(code, msg); }

// Natural order -> second-pass order: bin k = k1 + R k2 (k < M) to position
// k1 C + k2, i.e. the C x R matrix src[k2][k1] transposed, through a 32 x 32
// LDS tile so that reads and writes stay coalesced; the Nyquist bin stays at M.
template <typename T>
__global__ __launch_bounds__(256) void k_permute_bins(const T *__restrict__ src0, const T *__restrict__ src1,
                                                      T *__restrict__ dst, int R, int C, int M) {
  __shared__ T tile[32][33];
  const int ch = blockIdx.z;
  const T *src = ch ? src1 : src0;
  T *out = dst + (size_t)ch * (M + 1);
  const int k1_0 = blockIdx.x * 32, k2_0 = blockIdx.y * 32;
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int k2 = k2_0 + j, k1 = k1_0 + threadIdx.x;
    if (k1 < R && k2 < C) tile[j][threadIdx.x] = src[(size_t)k2 * R + k1];
  }
  __syncthreads();
  for (int j = threadIdx.y; j < 32; j += 8) {
    const int k1 = k1_0 + j, k2 = k2_0 + threadIdx.x;
    if (k1 < R && k2 < C) out[(size_t)k1 * C + k2] = tile[threadIdx.x][j];
  }
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) out[M] = src[M];
}

}  // namespace

extern "C" int aak_inner_product_noalloc(const double *hI, const double *hII, const double *hI_f,
                                         const double *hII_f, const double *ASDinv1, const double *ASDinv2,
                                         int length, double *like_out, void *stream) {
  FftArgs F{};
  int R = 0, C = 0;
  if (!fft_supported(length, &R, &C) || length > kMaxN)
    return fail(-2, "inner_product_noalloc: length must be a power of two in [2^8, 2^21], got " +
                        std::to_string(length));
  if (!hI || !hII || !hI_f || !hII_f || !ASDinv1 || !ASDinv2 || !like_out)
    return fail(-2, "inner_product_noalloc: null pointer");
  const uintptr_t mis = (uintptr_t)hI | (uintptr_t)hII | (uintptr_t)hI_f | (uintptr_t)hII_f;
  if (mis & 15) return fail(-2, "inner_product_noalloc: series and spectra must be 16-byte aligned");
  F.N = length;
  F.M = length / 2;
  F.R = R;
  F.C = C;
  hipStream_t s = reinterpret_cast<hipStream_t>(stream);
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail(-100, std::string("HIP: ") + hipGetErrorString(e) + " at hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(-2, "inner_product_noalloc: device index out of range");
  unsigned char *base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p)
        return fail(-100, std::string("HIP: ") + hipGetErrorString(e) + " at hipGetSymbolAddress(scratch)");
      g_scratch[dev] = static_cast<unsigned char *>(p);
    }
    base = g_scratch[dev];
    double2 *tw = reinterpret_cast<double2 *>(base + kOffTw);
    if (g_table_n[dev] != length) {
      // the tables of a new length: filled in a static host array (under the
      // lock) and copied; the call waits for the copy.  Later calls of this
      // length copy nothing and only enqueue kernels
      static double2 host[kBytesTw / sizeof(double2)];
      const size_t n = fft_table_size(F);
      if (n > kBytesTw / sizeof(double2)) return fail(-2, "inner_product_noalloc: twiddle tables too large");
      fft_fill_tables(F, host);
      e = hipMemcpyAsync(tw, host, sizeof(double2) * n, hipMemcpyHostToDevice, s);
      if (e == hipSuccess) e = hipStreamSynchronize(s);
      if (e != hipSuccess) {
        g_table_n[dev] = 0;
        return fail(-100, std::string("HIP: ") + hipGetErrorString(e) + " uploading the FFT tables");
      }
      g_table_n[dev] = length;
    }
    fft_bind_tables(&F, tw);
  }
  double2 *Y = reinterpret_cast<double2 *>(base + kOffY);
  double2 *D = reinterpret_cast<double2 *>(base + kOffD);
  double *W = reinterpret_cast<double *>(base + kOffW);
  double *partials = reinterpret_cast<double *>(base + kOffP);
  const size_t nb = (size_t)F.M + 1;
  F.Dp[0] = D;
  F.Dp[1] = D + nb;
  F.wp[0] = W;
  F.wp[1] = W + nb;
  const dim3 grid((R + 31) / 32, (C + 31) / 32, 2), block(32, 8);
  hipLaunchKernelGGL(k_permute_bins<double2>, grid, block, 0, s, reinterpret_cast<const double2 *>(hI_f),
                     reinterpret_cast<const double2 *>(hII_f), D, R, C, F.M);
  hipLaunchKernelGGL(k_permute_bins<double>, grid, block, 0, s, ASDinv1, ASDinv2, W, R, C, F.M);
  launch_fft_pass1(F, hI, hII, Y, s);
  launch_fft_pass2(F, Y, partials, nullptr, s);
  launch_dot_final(partials, fft_partials(F), like_out, nullptr, nullptr, s);
  e = hipGetLastError();
  if (e != hipSuccess) return fail(-100, std::string("HIP: ") + hipGetErrorString(e) + " launching the inner product");
  return 0;
}
