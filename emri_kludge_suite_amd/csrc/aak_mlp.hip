// aak_mlp.hip -- learned start point for the frequency map (KSParMap.cc:397-440):
// a 4-128-128-128-3 network with GELU activations, BF16 weights and FP32
// accumulation on the matrix cores (v_mfma_f32_16x16x32_bf16), with its
// training kernels (forward with saved pre-activations, backward chain,
// weight-gradient partials, Adam) and batched inference.
//
// The network maps the normalised features of a trajectory point
// (a, 1/sqrt(p), e, cos iota) to the standardised corrections
// (v_map - 1/sqrt(p), M_map/M - 1, s_map - a); the host turns them into the
// start point of a Newton polish of the map (aak_freqmap.hpp).
//
// Rounding points (the CPU gradient reference reproduces them): weight
// matrices in BF16, biases FP32; the inputs of every matrix-core product in
// BF16 (activations A1..A3, the output gradient dY and the back-propagated
// dZ3, dZ2); everything else FP32.  Layer 1 (K = 4) and the products with
// K = 3 or an output width of 3 on the weight-gradient side run on the VALU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/gpuAAK.h"
#include "aak_mlp.hpp"

namespace aak {
int set_last_error(int code, const std::string &msg);
}

using namespace aak;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int H = kMlpHidden;  // 128
constexpr int LDW = H + 8;     // LDS row stride (bf16): 272 B, conflict-free 16-byte fragment reads
constexpr int TILE = 64;       // rows per workgroup in the forward and backward chains
constexpr int CHUNK = 128;     // rows per workgroup in the weight-gradient partials
constexpr int kThreads = 256;

int fail(int code, const std::string &msg) { return set_last_error(code, msg); }

#define MLPCHK(x)                                                                                     \
  do {                                                                                                \
    hipError_t e_ = (x);                                                                              \
    if (e_ != hipSuccess) return fail(-100, std::string("HIP: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

__device__ __forceinline__ uint16_t bf16_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// GELU, tanh form, and its derivative
__device__ __forceinline__ float gelu(float z) {
  const float c = 0.7978845608028654f;
  return 0.5f * z * (1.0f + tanhf(c * (z + 0.044715f * z * z * z)));
}
__device__ __forceinline__ float gelu_grad(float z) {
  const float c = 0.7978845608028654f;
  const float t = tanhf(c * (z + 0.044715f * z * z * z));
  return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * c * (1.0f + 3.0f * 0.044715f * z * z);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t *p) {
  const uint4 v = *reinterpret_cast<const uint4 *>(p);
  return __builtin_bit_cast(bf16x8, v);
}

// Parameter offsets (floats) in the flat vector: W1[4][128] b1 W2[128][128]
// b2 W3[128][128] b3 W4[128][3] b4[3]; W[k][n] is input k to output n.
constexpr int oW1 = 0, oB1 = oW1 + 4 * H, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H * H,
              oW4 = oB3 + H, oB4 = oW4 + 3 * H, kNP = oB4 + 3;
static_assert(kNP == kMlpParams, "parameter count");

// Stage W (H x H, k-major bf16 in global) into LDS, transposed (row n holds
// input k contiguous) when `trans`, natural (row k holds n) otherwise.
__device__ void stage_hxh(uint16_t *lds, const uint16_t *W, bool trans) {
  for (int q = threadIdx.x; q < H * H / 8; q += kThreads) {
    const int k = q / (H / 8), n0 = (q % (H / 8)) * 8;
    const uint4 v = *reinterpret_cast<const uint4 *>(W + k * H + n0);
    const uint16_t *e = reinterpret_cast<const uint16_t *>(&v);
    if (trans) {
      for (int j = 0; j < 8; j++) lds[(n0 + j) * LDW + k] = e[j];
    } else {
      *reinterpret_cast<uint4 *>(lds + k * LDW + n0) = v;
    }
  }
}

// One 16-row x 128-column product per wave: acc[t] (columns 16t..16t+15,
// C layout: lane holds rows 4(l/16)+i of column l%16) += act[16 rows][K] x B,
// B fragments from `wl` (row = output column, K contiguous).
__device__ __forceinline__ void wave_gemm_128(floatx4 acc[8], const uint16_t *act, const uint16_t *wl, int K) {
  const int l = threadIdx.x & 63;
  for (int k0 = 0; k0 < K; k0 += 32) {
    const bf16x8 a = load8(act + (l & 15) * LDW + k0 + 8 * (l >> 4));
#pragma unroll
    for (int t = 0; t < 8; t++) {
      const bf16x8 b = load8(wl + (16 * t + (l & 15)) * LDW + k0 + 8 * (l >> 4));
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[t], 0, 0, 0);
    }
  }
}

// ---------------------------------------------------------------- forward
// Rows [blockIdx.x * 64, +64) of the batch; row r of the batch is row idx[r]
// of X/T (idx null: identity).  Inference writes Y only.  Training also
// writes the gathered inputs Xb, the pre-activations Z1..Z3 (FP32), dY and a
// per-workgroup partial of the squared error.
template <bool TRAIN>
__global__ __launch_bounds__(kThreads) void k_mlp_forward(const float *__restrict__ X, const float *__restrict__ T,
                                                          const int *__restrict__ idx, int n,
                                                          const float *__restrict__ P, const uint16_t *__restrict__ Pb,
                                                          float *__restrict__ Y, float *__restrict__ Xb,
                                                          float *__restrict__ Z, float *__restrict__ dY,
                                                          float *__restrict__ loss_part, float dscale) {
  __shared__ __align__(16) uint16_t wl[H * LDW];
  __shared__ __align__(16) uint16_t act[TILE * LDW];
  __shared__ float xs[TILE][4];
  __shared__ float red[kThreads / 64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * TILE;
  for (int q = threadIdx.x; q < TILE * 4; q += kThreads) {
    const int r = q >> 2, k = q & 3, gr = row0 + r;
    float x = 0.f;
    if (gr < n) x = X[(size_t)(idx ? idx[gr] : gr) * 4 + k];
    xs[r][k] = x;
    if (TRAIN && gr < n) Xb[(size_t)gr * 4 + k] = x;
  }
  __syncthreads();
  // layer 1 on the VALU: lane (row l%16 of the wave, columns 32 (l/16) + j)
  {
    const int r = 16 * w + (l & 15);
    const float x0 = xs[r][0], x1 = xs[r][1], x2 = xs[r][2], x3 = xs[r][3];
    for (int j = 0; j < 32; j++) {
      const int c = 32 * (l >> 4) + j;
      float z = P[oB1 + c];
      z = fmaf(x0, bf16_to_f32(Pb[oW1 + c]), z);
      z = fmaf(x1, bf16_to_f32(Pb[oW1 + H + c]), z);
      z = fmaf(x2, bf16_to_f32(Pb[oW1 + 2 * H + c]), z);
      z = fmaf(x3, bf16_to_f32(Pb[oW1 + 3 * H + c]), z);
      if (TRAIN && row0 + r < n) Z[(size_t)(row0 + r) * H + c] = z;
      act[r * LDW + c] = bf16_bits(gelu(z));
    }
  }
  // layers 2 and 3 on the matrix cores
  for (int layer = 0; layer < 2; layer++) {
    __syncthreads();
    stage_hxh(wl, Pb + (layer ? oW3 : oW2), true);
    __syncthreads();
    floatx4 acc[8];
    for (int t = 0; t < 8; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    wave_gemm_128(acc, act + 16 * w * LDW, wl, H);
    __syncthreads();
    const float *bias = P + (layer ? oB3 : oB2);
    float *Zl = TRAIN ? Z + (size_t)(layer + 1) * n * H : nullptr;
    for (int t = 0; t < 8; t++) {
      const int c = 16 * t + (l & 15);
      for (int i = 0; i < 4; i++) {
        const int r = 16 * w + 4 * (l >> 4) + i;
        const float z = acc[t][i] + bias[c];
        if (TRAIN && row0 + r < n) Zl[(size_t)(row0 + r) * H + c] = z;
        act[r * LDW + c] = bf16_bits(gelu(z));
      }
    }
  }
  // layer 4: output width 3, padded to 16 columns
  __syncthreads();
  for (int q = threadIdx.x; q < 16 * H; q += kThreads) {
    const int j = q / H, k = q % H;
    wl[j * LDW + k] = j < 3 ? Pb[oW4 + k * 3 + j] : (uint16_t)0;
  }
  __syncthreads();
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < H; k0 += 32) {
    const bf16x8 a = load8(act + (16 * w + (l & 15)) * LDW + k0 + 8 * (l >> 4));
    const bf16x8 b = load8(wl + (l & 15) * LDW + k0 + 8 * (l >> 4));
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  float se = 0.f;
  const int j = l & 15;
  if (j < 3) {
    for (int i = 0; i < 4; i++) {
      const int gr = row0 + 16 * w + 4 * (l >> 4) + i;
      if (gr >= n) continue;
      const float y = acc[i] + P[oB4 + j];
      if (Y) Y[(size_t)gr * 3 + j] = y;
      if (TRAIN) {
        const float d = y - T[(size_t)(idx ? idx[gr] : gr) * 3 + j];
        se += d * d;
        dY[(size_t)gr * 3 + j] = dscale * d;
      }
    }
  }
  if (TRAIN) {
    for (int o = 32; o > 0; o >>= 1) se += __shfl_xor(se, o);
    if (l == 0) red[w] = se;
    __syncthreads();
    if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
  }
}

// ------------------------------------------------------------ backward chain
// dA3 = dY W4^T, dZ3 = dA3 * gelu'(Z3), dA2 = dZ3 W3^T, dZ2, dA1 = dZ2 W2^T,
// dZ1; writes dZ1..dZ3 (FP32, [3][n][128]).
__global__ __launch_bounds__(kThreads) void k_mlp_backward(int n, const uint16_t *__restrict__ Pb,
                                                           const float *__restrict__ Z,
                                                           const float *__restrict__ dY, float *__restrict__ dZ) {
  __shared__ __align__(16) uint16_t wl[H * LDW];
  __shared__ __align__(16) uint16_t act[TILE * LDW];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * TILE;
  // dY (BF16, K padded to 32) as the A operand; W4 natural (row k, outputs j)
  for (int q = threadIdx.x; q < TILE * 32; q += kThreads) {
    const int r = q / 32, j = q % 32, gr = row0 + r;
    act[r * LDW + j] = (j < 3 && gr < n) ? bf16_bits(dY[(size_t)gr * 3 + j]) : (uint16_t)0;
  }
  for (int q = threadIdx.x; q < H * 32; q += kThreads) {
    const int k = q / 32, j = q % 32;
    wl[k * LDW + j] = j < 3 ? Pb[oW4 + k * 3 + j] : (uint16_t)0;
  }
  __syncthreads();
  floatx4 acc[8];
  for (int t = 0; t < 8; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
  wave_gemm_128(acc, act + 16 * w * LDW, wl, 32);
  for (int layer = 2; layer >= 0; layer--) {
    // acc = dA_(layer+1) in C layout: dZ = dA * gelu'(Z)
    const float *Zl = Z + (size_t)layer * n * H;
    float *dZl = dZ + (size_t)layer * n * H;
    __syncthreads();
    for (int t = 0; t < 8; t++) {
      const int c = 16 * t + (l & 15);
      for (int i = 0; i < 4; i++) {
        const int r = 16 * w + 4 * (l >> 4) + i, gr = row0 + r;
        float dz = 0.f;
        if (gr < n) {
          dz = acc[t][i] * gelu_grad(Zl[(size_t)gr * H + c]);
          dZl[(size_t)gr * H + c] = dz;
        }
        act[r * LDW + c] = bf16_bits(dz);
      }
    }
    if (layer == 0) break;
    stage_hxh(wl, Pb + (layer == 2 ? oW3 : oW2), false);
    __syncthreads();
    for (int t = 0; t < 8; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    wave_gemm_128(acc, act + 16 * w * LDW, wl, H);
  }
}

// ------------------------------------------------------- weight gradients
// Partial gradients over rows [blockIdx.x * 128, +128): dW2 = A1^T dZ2 and
// dW3 = A2^T dZ3 on the matrix cores (A = bf16(gelu(Z)), dZ rounded to BF16),
// dW1 = X^T dZ1, dW4 = A3^T dY and the bias sums on the VALU.
__global__ __launch_bounds__(kThreads) void k_mlp_wgrad(int n, const float *__restrict__ Xb,
                                                        const float *__restrict__ Z, const float *__restrict__ dZ,
                                                        const float *__restrict__ dY, float *__restrict__ part) {
  __shared__ __align__(16) uint16_t at[H * LDW];  // A^T: row = feature m, rows of the chunk contiguous
  __shared__ __align__(16) uint16_t dt[H * LDW];  // dZ^T
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * CHUNK;
  const int rows = min(CHUNK, n - row0);
  float *out = part + (size_t)blockIdx.x * kNP;
  for (int m = 0; m < 2; m++) {  // m = 0: dW2 (A1, dZ2), m = 1: dW3 (A2, dZ3)
    const float *Zs = Z + (size_t)m * n * H;
    const float *dZs = dZ + (size_t)(m + 1) * n * H;
    __syncthreads();
    for (int q = threadIdx.x; q < CHUNK * H; q += kThreads) {
      const int r = q / H, c = q % H;
      uint16_t a = 0, d = 0;
      if (r < rows) {
        a = bf16_bits(gelu(Zs[(size_t)(row0 + r) * H + c]));
        d = bf16_bits(dZs[(size_t)(row0 + r) * H + c]);
      }
      at[c * LDW + r] = a;
      dt[c * LDW + r] = d;
    }
    __syncthreads();
    // 8 x 8 output tiles of 16 x 16; wave w takes tile rows 2w, 2w + 1
    for (int tr = 2 * w; tr < 2 * w + 2; tr++) {
      floatx4 acc[8];
      for (int t = 0; t < 8; t++) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
      wave_gemm_128(acc, at + 16 * tr * LDW, dt, CHUNK);
      float *dW = out + (m ? oW3 : oW2);
      for (int t = 0; t < 8; t++)
        for (int i = 0; i < 4; i++) dW[(16 * tr + 4 * (l >> 4) + i) * H + 16 * t + (l & 15)] = acc[t][i];
    }
  }
  // dW1[k][c] = sum_r X[r][k] dZ1[r][c]; db_l[c] = sum_r dZ_l[r][c]
  for (int c = threadIdx.x; c < H; c += kThreads) {
    float g0 = 0.f, g1 = 0.f, g2 = 0.f, g3 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
    for (int r = 0; r < rows; r++) {
      const size_t gr = row0 + r;
      const float d1 = dZ[gr * H + c];
      g0 = fmaf(Xb[gr * 4 + 0], d1, g0);
      g1 = fmaf(Xb[gr * 4 + 1], d1, g1);
      g2 = fmaf(Xb[gr * 4 + 2], d1, g2);
      g3 = fmaf(Xb[gr * 4 + 3], d1, g3);
      s1 += d1;
      s2 += dZ[(size_t)n * H + gr * H + c];
      s3 += dZ[(size_t)2 * n * H + gr * H + c];
    }
    out[oW1 + c] = g0;
    out[oW1 + H + c] = g1;
    out[oW1 + 2 * H + c] = g2;
    out[oW1 + 3 * H + c] = g3;
    out[oB1 + c] = s1;
    out[oB2 + c] = s2;
    out[oB3 + c] = s3;
  }
  // dW4[k][j] = sum_r A3[r][k] dY[r][j] (A3 = bf16(gelu(Z3))); db4
  for (int q = threadIdx.x; q < 3 * H + 3; q += kThreads) {
    float g = 0.f;
    if (q < 3 * H) {
      const int k = q / 3, j = q % 3;
      for (int r = 0; r < rows; r++) {
        const size_t gr = row0 + r;
        g = fmaf(bf16_to_f32(bf16_bits(gelu(Z[(size_t)2 * n * H + gr * H + k]))), dY[gr * 3 + j], g);
      }
      out[oW4 + q] = g;
    } else {
      const int j = q - 3 * H;
      for (int r = 0; r < rows; r++) g += dY[(size_t)(row0 + r) * 3 + j];
      out[oB4 + j] = g;
    }
  }
}

// ----------------------------------------------------------------- Adam
// g = sum of the chunk partials; with step > 0 an Adam step on the FP32
// master copy, whose BF16 copy is refreshed; with step == 0 only g is written.
__global__ __launch_bounds__(kThreads) void k_mlp_adam(int nchunks, const float *__restrict__ part,
                                                       float *__restrict__ grad, float *__restrict__ P,
                                                       uint16_t *__restrict__ Pb, float *__restrict__ m1,
                                                       float *__restrict__ m2, int step, float lr, float b1,
                                                       float b2, float eps) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= kNP) return;
  float g = 0.f;
  for (int c = 0; c < nchunks; c++) g += part[(size_t)c * kNP + p];
  if (grad) grad[p] = g;
  if (step <= 0) return;
  const float a = b1 * m1[p] + (1.f - b1) * g, v = b2 * m2[p] + (1.f - b2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float ah = a / (1.f - powf(b1, (float)step)), vh = v / (1.f - powf(b2, (float)step));
  const float x = P[p] - lr * ah / (sqrtf(vh) + eps);
  P[p] = x;
  Pb[p] = bf16_bits(x);
}

__global__ void k_mlp_round(const float *__restrict__ P, uint16_t *__restrict__ Pb) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p < kNP) Pb[p] = bf16_bits(P[p]);
}

}  // namespace

// ------------------------------------------------------------------ host
struct aak_mlp_s {
  float *P = nullptr, *m1 = nullptr, *m2 = nullptr, *grad = nullptr;
  uint16_t *Pb = nullptr;
  int step = 0;
  float norm[kMlpNorm] = {};  // feature centres/scales and target means/scales
  hipStream_t s = nullptr;
  int dev = 0;  // device the model lives on (callers may be other host threads)
  // training / gradient scratch (grown on demand)
  int cap = 0;
  float *X = nullptr, *T = nullptr, *Xb = nullptr, *Z = nullptr, *dZ = nullptr, *dY = nullptr, *loss = nullptr,
        *part = nullptr;
  int *idx = nullptr;
  int data_n = 0;
  // inference scratch
  std::mutex mu;
  int icap = 0;
  float *iX = nullptr, *iY = nullptr;
  std::vector<float> hx, hy;
};


namespace aak {

static void free_scratch(aak_mlp_s *m) {
  void *ps[] = {m->X, m->T, m->Xb, m->Z, m->dZ, m->dY, m->loss, m->part, m->idx};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  m->X = m->T = m->Xb = m->Z = m->dZ = m->dY = m->loss = m->part = nullptr;
  m->idx = nullptr;
  m->cap = 0;
  m->data_n = 0;
}

// Scratch for batches of up to B rows over a data set of `data` rows.
static int grow_scratch(aak_mlp_s *m, int B, int data) {
  if (B > m->cap) {
    void *ps[] = {m->Xb, m->Z, m->dZ, m->dY, m->loss, m->part};
    for (void *p : ps)
      if (p) (void)hipFree(p);
    const int tiles = (B + TILE - 1) / TILE, chunks = (B + CHUNK - 1) / CHUNK;
    MLPCHK(hipMalloc(&m->Xb, sizeof(float) * 4 * (size_t)B));
    MLPCHK(hipMalloc(&m->Z, sizeof(float) * 3 * (size_t)B * H));
    MLPCHK(hipMalloc(&m->dZ, sizeof(float) * 3 * (size_t)B * H));
    MLPCHK(hipMalloc(&m->dY, sizeof(float) * 3 * (size_t)B));
    MLPCHK(hipMalloc(&m->loss, sizeof(float) * tiles));
    MLPCHK(hipMalloc(&m->part, sizeof(float) * (size_t)chunks * kNP));
    m->cap = B;
  }
  if (data > m->data_n) {
    if (m->X) (void)hipFree(m->X);
    if (m->T) (void)hipFree(m->T);
    if (m->idx) (void)hipFree(m->idx);
    MLPCHK(hipMalloc(&m->X, sizeof(float) * 4 * (size_t)data));
    MLPCHK(hipMalloc(&m->T, sizeof(float) * 3 * (size_t)data));
    MLPCHK(hipMalloc(&m->idx, sizeof(int) * (size_t)data));
    m->data_n = data;
  }
  return 0;
}

// One forward + backward over rows idx[off .. off+B) of the uploaded data;
// with step > 0 an Adam update.  loss_sum (optional) gets the squared error.
static int train_step(aak_mlp_s *m, int off, int B, float lr, bool update, double *loss_sum) {
  const int tiles = (B + TILE - 1) / TILE, chunks = (B + CHUNK - 1) / CHUNK;
  const float dscale = 2.f / (3.f * (float)B);
  hipLaunchKernelGGL(k_mlp_forward<true>, dim3(tiles), dim3(kThreads), 0, m->s, m->X, m->T, m->idx + off, B, m->P,
                     m->Pb, nullptr, m->Xb, m->Z, m->dY, m->loss, dscale);
  hipLaunchKernelGGL(k_mlp_backward, dim3(tiles), dim3(kThreads), 0, m->s, B, m->Pb, m->Z, m->dY, m->dZ);
  hipLaunchKernelGGL(k_mlp_wgrad, dim3(chunks), dim3(kThreads), 0, m->s, B, m->Xb, m->Z, m->dZ, m->dY, m->part);
  if (update) m->step++;
  hipLaunchKernelGGL(k_mlp_adam, dim3((kNP + kThreads - 1) / kThreads), dim3(kThreads), 0, m->s, chunks, m->part,
                     m->grad, m->P, m->Pb, m->m1, m->m2, update ? m->step : 0, lr, 0.9f, 0.999f, 1e-8f);
  MLPCHK(hipGetLastError());
  if (loss_sum) {
    std::vector<float> lp(tiles);
    MLPCHK(hipMemcpyAsync(lp.data(), m->loss, sizeof(float) * tiles, hipMemcpyDeviceToHost, m->s));
    MLPCHK(hipStreamSynchronize(m->s));
    double s = 0.0;
    for (float x : lp) s += x;
    *loss_sum = s;
  }
  return 0;
}

// Normalised features of (a, p, e, iota) and the corrections of a map
// (v_map, M_map/M, s_map) -- see the file comment.
static void features(const aak_mlp_s *m, const double *pt, float *x) {
  const double a = pt[0], p = pt[1], e = pt[2], ci = cos(pt[3]);
  const double f[4] = {a, 1.0 / sqrt(p), e, ci};
  for (int k = 0; k < 4; k++) x[k] = (float)((f[k] - m->norm[k]) / m->norm[4 + k]);
}

int mlp_infer(aak_mlp_s *m, const double *pts, int n, double *maps) {
  if (n <= 0) return 0;
  std::lock_guard<std::mutex> lk(m->mu);
  MLPCHK(hipSetDevice(m->dev));
  if (n > m->icap) {
    if (m->iX) (void)hipFree(m->iX);
    if (m->iY) (void)hipFree(m->iY);
    m->iX = m->iY = nullptr;
    m->icap = 0;
    MLPCHK(hipMalloc(&m->iX, sizeof(float) * 4 * (size_t)n));
    MLPCHK(hipMalloc(&m->iY, sizeof(float) * 3 * (size_t)n));
    m->icap = n;
  }
  m->hx.resize(4 * (size_t)n);
  m->hy.resize(3 * (size_t)n);
  for (int i = 0; i < n; i++) features(m, pts + 4 * (size_t)i, m->hx.data() + 4 * (size_t)i);
  MLPCHK(hipMemcpyAsync(m->iX, m->hx.data(), sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, m->s));
  hipLaunchKernelGGL(k_mlp_forward<false>, dim3((n + TILE - 1) / TILE), dim3(kThreads), 0, m->s, m->iX, nullptr,
                     nullptr, n, m->P, m->Pb, m->iY, nullptr, nullptr, nullptr, nullptr, 0.f);
  MLPCHK(hipGetLastError());
  MLPCHK(hipMemcpyAsync(m->hy.data(), m->iY, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, m->s));
  MLPCHK(hipStreamSynchronize(m->s));
  for (int i = 0; i < n; i++) {
    const double *pt = pts + 4 * (size_t)i;
    const float *y = m->hy.data() + 3 * (size_t)i;
    const double c[3] = {m->norm[8] + m->norm[11] * (double)y[0], m->norm[9] + m->norm[12] * (double)y[1],
                         m->norm[10] + m->norm[13] * (double)y[2]};
    maps[3 * (size_t)i + 0] = 1.0 / sqrt(pt[1]) + c[0];
    maps[3 * (size_t)i + 1] = 1.0 + c[1];
    maps[3 * (size_t)i + 2] = pt[0] + c[2];
  }
  return 0;
}

}  // namespace aak

// ------------------------------------------------------------------ C ABI
extern "C" {

int aak_mlp_create(int seed, aak_mlp *out) {
  if (!out) return fail(-2, "null model pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-100, "no HIP device");
  aak_mlp_s *m = new aak_mlp_s();
  if (hipGetDevice(&m->dev) != hipSuccess) m->dev = 0;
  auto bail = [&](int rc) {
    aak_mlp_destroy(m);
    return rc;
  };
  if (hipStreamCreateWithFlags(&m->s, hipStreamNonBlocking) != hipSuccess) return bail(fail(-100, "stream"));
  float *bufs[4];
  for (int k = 0; k < 4; k++)
    if (hipMalloc(&bufs[k], sizeof(float) * kNP) != hipSuccess) {
      for (int j = 0; j < k; j++) (void)hipFree(bufs[j]);
      return bail(fail(-100, "device allocation of the parameters failed"));
    }
  m->P = bufs[0];
  m->m1 = bufs[1];
  m->m2 = bufs[2];
  m->grad = bufs[3];
  if (hipMalloc(&m->Pb, sizeof(uint16_t) * kNP) != hipSuccess) return bail(fail(-100, "device allocation failed"));
  // He-style initialisation; identity normalisation until the trainer sets it
  std::mt19937_64 rng((uint64_t)seed);
  std::vector<float> p(kNP, 0.f);
  auto fill = [&](int off, int fan_in, int count) {
    std::normal_distribution<float> d(0.f, sqrtf(2.f / (float)fan_in));
    for (int i = 0; i < count; i++) p[off + i] = d(rng);
  };
  fill(oW1, 4, 4 * H);
  fill(oW2, H, H * H);
  fill(oW3, H, H * H);
  fill(oW4, H, 3 * H);
  for (int k = 0; k < 4; k++) m->norm[4 + k] = 1.f;
  for (int k = 0; k < 3; k++) m->norm[11 + k] = 1.f;
  const int rc = aak_mlp_set_params(m, p.data(), nullptr);
  if (rc) return bail(rc);
  std::vector<float> zero(kNP, 0.f);
  if (hipMemcpy(m->m1, zero.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->m2, zero.data(), sizeof(float) * kNP, hipMemcpyHostToDevice) != hipSuccess)
    return bail(fail(-100, "upload failed"));
  *out = m;
  return 0;
}

int aak_mlp_destroy(aak_mlp m) {
  if (!m) return fail(-1, "null model");
  if (m->s) (void)hipStreamSynchronize(m->s);
  free_scratch(m);
  void *ps[] = {m->P, m->m1, m->m2, m->grad, m->Pb, m->iX, m->iY};
  for (void *p : ps)
    if (p) (void)hipFree(p);
  if (m->s) (void)hipStreamDestroy(m->s);
  delete m;
  return 0;
}

int aak_mlp_set_params(aak_mlp m, const float *params, const float *norm) {
  if (!m || !params) return fail(-2, "null argument");
  MLPCHK(hipMemcpyAsync(m->P, params, sizeof(float) * kNP, hipMemcpyHostToDevice, m->s));
  hipLaunchKernelGGL(k_mlp_round, dim3((kNP + kThreads - 1) / kThreads), dim3(kThreads), 0, m->s, m->P, m->Pb);
  MLPCHK(hipGetLastError());
  MLPCHK(hipStreamSynchronize(m->s));
  if (norm) std::memcpy(m->norm, norm, sizeof(m->norm));
  m->step = 0;
  return 0;
}

int aak_mlp_get_params(aak_mlp m, float *params, float *norm) {
  if (!m) return fail(-1, "null model");
  if (params) {
    MLPCHK(hipMemcpyAsync(params, m->P, sizeof(float) * kNP, hipMemcpyDeviceToHost, m->s));
    MLPCHK(hipStreamSynchronize(m->s));
  }
  if (norm) std::memcpy(norm, m->norm, sizeof(m->norm));
  return kNP;
}

int aak_mlp_forward(aak_mlp m, const float *X, int n, float *Y) {
  if (!m || !X || !Y || n < 0) return fail(-2, "bad argument");
  if (n == 0) return 0;
  std::lock_guard<std::mutex> lk(m->mu);
  MLPCHK(hipSetDevice(m->dev));
  if (n > m->icap) {
    if (m->iX) (void)hipFree(m->iX);
    if (m->iY) (void)hipFree(m->iY);
    m->iX = m->iY = nullptr;
    m->icap = 0;
    MLPCHK(hipMalloc(&m->iX, sizeof(float) * 4 * (size_t)n));
    MLPCHK(hipMalloc(&m->iY, sizeof(float) * 3 * (size_t)n));
    m->icap = n;
  }
  MLPCHK(hipMemcpyAsync(m->iX, X, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, m->s));
  hipLaunchKernelGGL(k_mlp_forward<false>, dim3((n + TILE - 1) / TILE), dim3(kThreads), 0, m->s, m->iX, nullptr,
                     nullptr, n, m->P, m->Pb, m->iY, nullptr, nullptr, nullptr, nullptr, 0.f);
  MLPCHK(hipGetLastError());
  MLPCHK(hipMemcpyAsync(Y, m->iY, sizeof(float) * 3 * (size_t)n, hipMemcpyDeviceToHost, m->s));
  MLPCHK(hipStreamSynchronize(m->s));
  return 0;
}

int aak_mlp_gradient(aak_mlp m, const float *X, const float *T, int n, float *grad, double *loss) {
  if (!m || !X || !T || !grad || n <= 0) return fail(-2, "bad argument");
  MLPCHK(hipSetDevice(m->dev));
  int rc = grow_scratch(m, n, n);
  if (rc) return rc;
  std::vector<int> id(n);
  for (int i = 0; i < n; i++) id[i] = i;
  MLPCHK(hipMemcpyAsync(m->X, X, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, m->s));
  MLPCHK(hipMemcpyAsync(m->T, T, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, m->s));
  MLPCHK(hipMemcpyAsync(m->idx, id.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->s));
  double ls = 0.0;
  rc = train_step(m, 0, n, 0.f, false, &ls);
  if (rc) return rc;
  MLPCHK(hipMemcpyAsync(grad, m->grad, sizeof(float) * kNP, hipMemcpyDeviceToHost, m->s));
  MLPCHK(hipStreamSynchronize(m->s));
  if (loss) *loss = ls / (3.0 * n);
  return 0;
}

int aak_mlp_train(aak_mlp m, const float *X, const float *T, int n, int epochs, int batch, float lr0, float lr1,
                  int seed, double *loss_per_epoch) {
  if (!m || !X || !T || n <= 0 || epochs <= 0 || batch <= 0 || batch > n) return fail(-2, "bad argument");
  MLPCHK(hipSetDevice(m->dev));
  int rc = grow_scratch(m, batch, n);
  if (rc) return rc;
  MLPCHK(hipMemcpyAsync(m->X, X, sizeof(float) * 4 * (size_t)n, hipMemcpyHostToDevice, m->s));
  MLPCHK(hipMemcpyAsync(m->T, T, sizeof(float) * 3 * (size_t)n, hipMemcpyHostToDevice, m->s));
  std::mt19937_64 rng((uint64_t)seed);
  std::vector<int> id(n);
  for (int i = 0; i < n; i++) id[i] = i;
  const int steps_per_epoch = n / batch;
  const long long total = (long long)epochs * steps_per_epoch;
  long long done = 0;
  for (int ep = 0; ep < epochs; ep++) {
    std::shuffle(id.begin(), id.end(), rng);
    MLPCHK(hipMemcpyAsync(m->idx, id.data(), sizeof(int) * (size_t)n, hipMemcpyHostToDevice, m->s));
    // the index upload must finish before `id` is reshuffled
    MLPCHK(hipStreamSynchronize(m->s));
    double ep_loss = 0.0;
    for (int st = 0; st < steps_per_epoch; st++, done++) {
      // cosine decay from lr0 to lr1
      const double f = 0.5 * (1.0 + cos(M_PI * (double)done / (double)total));
      const float lr = (float)(lr1 + (lr0 - lr1) * f);
      double ls = 0.0;
      rc = train_step(m, st * batch, batch, lr, true, (st == steps_per_epoch - 1) ? &ls : nullptr);
      if (rc) return rc;
      ep_loss = ls / (3.0 * batch);
    }
    if (loss_per_epoch) loss_per_epoch[ep] = ep_loss;
  }
  MLPCHK(hipStreamSynchronize(m->s));
  return 0;
}

int aak_mlp_set_norm(aak_mlp m, const float *norm) {
  if (!m || !norm) return fail(-2, "null argument");
  std::memcpy(m->norm, norm, sizeof(m->norm));
  return 0;
}

int aak_mlp_predict_maps(aak_mlp m, const double *pts, int n, double *maps) {
  if (!m || (n > 0 && (!pts || !maps)) || n < 0) return fail(-2, "bad argument");
  return mlp_infer(m, pts, n, maps);
}

}  // extern "C"
