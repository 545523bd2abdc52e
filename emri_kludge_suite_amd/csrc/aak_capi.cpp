// aak_capi.cpp -- the C ABI (include/gpuAAK.h, include/aak_testing.h).
//
// One handle owns a HIP stream, the device buffers of one waveform slot, the
// rocFFT real-forward plan (both channels in one batched call) and the
// weighted-data buffers.  gpu_gen_AAK runs the host stage, uploads the spline
// table and enqueues rates -> exact folds -> mode sum; Likelihood enqueues the
// rFFT and the fused weighted reduction and returns two doubles.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/aak_testing.h"
#include "../../include/gpuAAK.h"
#include "aak_dense.hpp"
#include "aak_geodesic.hpp"
#include "aak_host.hpp"

using namespace aak;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess) return fail(-100, std::string("HIP: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

std::once_flag g_fft_once;

}  // namespace

struct aak_handle_s {
  WaveConfig cfg;
  int N = 0, nbins = 0;
  hipStream_t stream = nullptr;
  // waveform slot
  double *d_knots = nullptr;
  size_t knots_cap = 0;
  double *d_dense = nullptr;  // e, v, M, S, rate[3], phase[3]: 10 x (N+1)
  double *d_h = nullptr;      // [2][N+2]
  double *d_t = nullptr;      // [N+2]
  int *d_flags = nullptr;
  // fold work
  FoldWork fw{};
  int fold_nblk_cap = 0;
  void *d_foldmem = nullptr;
  // likelihood
  double2 *d_hf = nullptr;    // [2][nbins]
  double2 *d_data = nullptr;  // [2][nbins]
  double *d_w = nullptr;      // [2][nbins]
  double *d_partials = nullptr, *d_like = nullptr;
  rocfft_plan plan = nullptr;
  rocfft_execution_info exec = nullptr;
  void *d_work = nullptr;
  int have_data = 0;
  int state = 0;  // 0 empty, 1 time domain, 2 weighted spectrum
  int info[10] = {0};
  // stage timing on the handle's stream: rates, folds, mode sum, rFFT, reduction
  int profiling = 0;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double stage_ms[5] = {0, 0, 0, 0, 0};
  double host_ms = 0.0;
  void mark(int k) {
    if (profiling) (void)hipEventRecord(ev[k], stream);
  }
  DenseBuffers bufs() const {
    DenseBuffers B;
    const size_t L = (size_t)N + 1;
    B.knots = d_knots;
    B.e = d_dense;
    B.v = d_dense + L;
    B.M = d_dense + 2 * L;
    B.S = d_dense + 3 * L;
    for (int p = 0; p < 3; p++) {
      B.rate[p] = d_dense + (4 + p) * L;
      B.phase[p] = d_dense + (7 + p) * L;
    }
    B.h = d_h;
    B.t = d_t;
    B.flags = d_flags;
    return B;
  }
};

static int alloc_handle(aak_handle h) {
  const int N = h->N;
  HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIPCHK(hipMalloc(&h->d_dense, sizeof(double) * 10 * ((size_t)N + 1)));
  HIPCHK(hipMalloc(&h->d_h, sizeof(double) * 2 * ((size_t)N + 2)));
  HIPCHK(hipMalloc(&h->d_t, sizeof(double) * ((size_t)N + 2)));
  HIPCHK(hipMalloc(&h->d_flags, sizeof(int) * 4));
  const int nblk = (N + kFoldBlock - 1) / kFoldBlock + 1;
  h->fold_nblk_cap = nblk;
  size_t bytes = 0;
  auto carve = [&](size_t n) {
    size_t off = bytes;
    bytes += (n + 255) & ~size_t(255);
    return off;
  };
  const size_t o_bsum = carve(sizeof(double) * 3 * nblk), o_bmin = carve(sizeof(double) * 3 * nblk);
  const size_t o_bmax = carve(sizeof(double) * 3 * nblk), o_xs = carve(sizeof(double) * 3 * nblk);
  const size_t o_E = carve(sizeof(double) * 3), o_ni = carve(sizeof(int) * 3 * nblk), o_nt = carve(sizeof(int) * 3 * nblk);
  const size_t o_items = carve(sizeof(FoldItem) * 3 * (size_t)nblk * kFoldMaxItems);
  const size_t o_ties = carve(sizeof(FoldTie) * 3 * (size_t)nblk * kFoldMaxTies);
  const size_t o_st = carve(sizeof(int) * 3);
  HIPCHK(hipMalloc(&h->d_foldmem, bytes));
  char *base = (char *)h->d_foldmem;
  h->fw.bsum = (double *)(base + o_bsum);
  h->fw.bmin = (double *)(base + o_bmin);
  h->fw.bmax = (double *)(base + o_bmax);
  h->fw.xstart = (double *)(base + o_xs);
  h->fw.E = (double *)(base + o_E);
  h->fw.nitems = (int *)(base + o_ni);
  h->fw.nties = (int *)(base + o_nt);
  h->fw.items = (FoldItem *)(base + o_items);
  h->fw.ties = (FoldTie *)(base + o_ties);
  h->fw.status = (int *)(base + o_st);
  const int nb = h->nbins;
  HIPCHK(hipMalloc(&h->d_hf, sizeof(double2) * 2 * (size_t)nb));
  HIPCHK(hipMalloc(&h->d_data, sizeof(double2) * 2 * (size_t)nb));
  HIPCHK(hipMalloc(&h->d_w, sizeof(double) * 2 * (size_t)nb));
  HIPCHK(hipMalloc(&h->d_partials, sizeof(double) * 2 * 2048));
  HIPCHK(hipMalloc(&h->d_like, sizeof(double) * 2));
  // rocFFT: two real sequences of length N (distance N+2) -> two half spectra
  std::call_once(g_fft_once, []() { rocfft_setup(); });
  rocfft_plan_description desc = nullptr;
  if (rocfft_plan_description_create(&desc) != rocfft_status_success) return fail(-101, "rocfft description");
  const size_t in_stride = 1, out_stride = 1;
  const size_t in_dist = (size_t)N + 2, out_dist = (size_t)nb;
  if (rocfft_plan_description_set_data_layout(desc, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved,
                                              nullptr, nullptr, 1, &in_stride, in_dist, 1, &out_stride,
                                              out_dist) != rocfft_status_success)
    return fail(-101, "rocfft layout");
  const size_t len = (size_t)N;
  if (rocfft_plan_create(&h->plan, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                         rocfft_precision_double, 1, &len, 2, desc) != rocfft_status_success)
    return fail(-101, "rocfft plan");
  rocfft_plan_description_destroy(desc);
  size_t wsize = 0;
  rocfft_plan_get_work_buffer_size(h->plan, &wsize);
  rocfft_execution_info_create(&h->exec);
  if (wsize) {
    HIPCHK(hipMalloc(&h->d_work, wsize));
    rocfft_execution_info_set_work_buffer(h->exec, h->d_work, wsize);
  }
  rocfft_execution_info_set_stream(h->exec, h->stream);
  return 0;
}

static void free_handle(aak_handle h) {
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  if (h->plan) rocfft_plan_destroy(h->plan);
  if (h->exec) rocfft_execution_info_destroy(h->exec);
  void *ptrs[] = {h->d_knots, h->d_dense, h->d_h, h->d_t, h->d_flags, h->d_foldmem, h->d_hf,
                  h->d_data, h->d_w, h->d_partials, h->d_like, h->d_work};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (hipEvent_t &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

extern "C" {

const char *aak_last_error(void) { return g_err.c_str(); }

double aak_OmegaPhi(double v, double e, double cosiota, double s, double M) { return omega_phi(v, e, cosiota, s, M); }

int aak_create(double T_fit, int init_length, int length, double init_dt, double dt, int LISA, int backint,
               aak_handle *out) {
  (void)init_length;
  (void)init_dt;
  if (!out) return fail(-1, "null handle pointer");
  *out = nullptr;
  if (length < 16 || (length & 1)) return fail(-2, "length must be an even number >= 16");
  if (!(dt > 0.0)) return fail(-2, "dt must be positive");
  if (!(T_fit > 0.0)) return fail(-2, "T_fit must be positive");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(-3, "no HIP device available");
  aak_handle h = new aak_handle_s();
  h->cfg.T_fit = T_fit;
  h->cfg.length = length;
  h->cfg.dt = dt;
  h->cfg.LISA = LISA ? 1 : 0;
  h->cfg.backint = backint ? 1 : 0;
  h->N = length;
  h->nbins = length / 2 + 1;
  const int rc = alloc_handle(h);
  if (rc) {
    std::string keep = g_err;
    free_handle(h);
    delete h;
    g_err = keep;
    return rc;
  }
  *out = h;
  return 0;
}

int aak_destroy(aak_handle h) {
  if (!h) return fail(-1, "null handle");
  free_handle(h);
  delete h;
  return 0;
}

int aak_set_nmax(aak_handle h, int n_max) {
  if (!h) return fail(-1, "null handle");
  if (n_max > 200) return fail(-2, "n_max above 200");
  h->cfg.n_max = n_max > 0 ? n_max : 0;
  return 0;
}

int aak_input_data(aak_handle h, const double *hI_f, const double *hII_f, const double *ASDinv1,
                   const double *ASDinv2, int len) {
  if (!h) return fail(-1, "null handle");
  if (len != h->nbins) return fail(-2, "input_data: len must equal length/2+1");
  const size_t nb = (size_t)h->nbins;
  HIPCHK(hipMemcpyAsync(h->d_data, hI_f, sizeof(double2) * nb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->d_data + nb, hII_f, sizeof(double2) * nb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->d_w, ASDinv1, sizeof(double) * nb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemcpyAsync(h->d_w + nb, ASDinv2, sizeof(double) * nb, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  h->have_data = 1;
  return 0;
}

int aak_gpu_gen_AAK(aak_handle h, double iota, double s, double p, double e, double M, double mu, double gamma,
                    double psi, double alph, double theta_S, double phi_S, double theta_K, double phi_K, double D) {
  if (!h) return fail(-1, "null handle");
  SourceParams sp{iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D};
  WaveformPlan plan;
  std::string err;
  auto t0 = std::chrono::steady_clock::now();
  int rc = plan_waveform(h->cfg, sp, &plan, &err);
  h->host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc < 0) return fail(rc, err);
  const DenseParams &P = plan.P;
  const size_t need = plan.packed.size();
  if (need > h->knots_cap) {
    if (h->d_knots) HIPCHK(hipFree(h->d_knots));
    h->d_knots = nullptr;
    const size_t cap = need + need / 4;
    HIPCHK(hipMalloc(&h->d_knots, sizeof(double) * cap));
    h->knots_cap = cap;
  }
  HIPCHK(hipMemcpyAsync(h->d_knots, plan.packed.data(), sizeof(double) * need, hipMemcpyHostToDevice, h->stream));
  HIPCHK(hipMemsetAsync(h->d_flags, 0, sizeof(int) * 4, h->stream));
  const DenseBuffers B = h->bufs();
  h->mark(0);
  launch_rates(P, B, h->stream);
  h->mark(1);
  int flags[4] = {0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(flags, h->d_flags, sizeof(int) * 4, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (flags[0] > 0) launch_phidot_guard(P, B, h->stream);
  int fallbacks = 0;
  const int nfwd = P.last_fold - P.i0 + 1;
  for (int pass = 0; pass < 2; pass++) {
    const int dir = pass == 0 ? 1 : -1;
    const int nsteps = pass == 0 ? nfwd : P.i0;
    if (nsteps <= 0) continue;
    if (fold_exact(P, B, h->fw, h->fold_nblk_cap, dir, nsteps, h->stream)) return fail(-7, "fold block capacity");
    int st[3];
    HIPCHK(hipMemcpyAsync(st, h->fw.status, sizeof(int) * 3, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    fallbacks += st[0] + st[1] + st[2];
  }
  launch_fold_fill(P, B, h->stream);
  h->mark(2);
  launch_modesum(P, B, h->stream);
  h->mark(3);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->profiling) {
    float ms = 0.f;
    for (int k = 0; k < 3; k++) {
      HIPCHK(hipEventElapsedTime(&ms, h->ev[k], h->ev[k + 1]));
      h->stage_ms[k] = ms;
    }
  }
  h->state = 1;
  int *I = h->info;
  I[0] = plan.traj.steps;
  I[1] = P.nmodes;
  I[2] = P.ip;
  I[3] = P.ib;
  I[4] = P.K;
  I[5] = P.i0;
  I[6] = fallbacks;
  I[7] = flags[0];
  I[8] = plan.knots.j0;
  I[9] = plan.knots.plunged;
  return 0;
}

int aak_likelihood(aak_handle h, double *like_out) {
  if (!h) return fail(-1, "null handle");
  if (!h->have_data) return fail(-2, "Likelihood before input_data");
  if (h->state != 1) return fail(-2, "Likelihood needs a waveform from gpu_gen_AAK");
  void *in[1] = {h->d_h};
  void *outb[1] = {h->d_hf};
  h->mark(3);
  if (rocfft_execute(h->plan, in, outb, h->exec) != rocfft_status_success) return fail(-101, "rocfft execute");
  h->mark(4);
  launch_weighted_dot(h->d_hf, h->d_data, h->d_w, h->nbins, h->d_hf, h->d_partials, h->d_like, h->stream);
  h->mark(5);
  HIPCHK(hipMemcpyAsync(like_out, h->d_like, sizeof(double) * 2, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  if (h->profiling) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, h->ev[3], h->ev[4]));
    h->stage_ms[3] = ms;
    HIPCHK(hipEventElapsedTime(&ms, h->ev[4], h->ev[5]));
    h->stage_ms[4] = ms;
  }
  h->state = 2;
  return 0;
}

int aak_get_waveform(aak_handle h, double *t, double *hI, double *hII) {
  if (!h) return fail(-1, "null handle");
  if (h->state == 0) return fail(-2, "GetWaveform before gpu_gen_AAK");
  const size_t n = (size_t)h->N + 2;
  HIPCHK(hipMemcpyAsync(t, h->d_t, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  if (h->state == 2) {
    HIPCHK(hipMemcpyAsync(hI, h->d_hf, sizeof(double2) * h->nbins, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hII, h->d_hf + h->nbins, sizeof(double2) * h->nbins, hipMemcpyDeviceToHost, h->stream));
  } else {
    HIPCHK(hipMemcpyAsync(hI, h->d_h, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(hII, h->d_h + n, sizeof(double) * n, hipMemcpyDeviceToHost, h->stream));
  }
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

int aak_set_profiling(aak_handle h, int on) {
  if (!h) return fail(-1, "null handle");
  if (on && !h->ev[0])
    for (hipEvent_t &e : h->ev) HIPCHK(hipEventCreate(&e));
  h->profiling = on ? 1 : 0;
  return 0;
}

int aak_get_timings(aak_handle h, double *ms, int n) {
  if (!h) return fail(-1, "null handle");
  const double v[6] = {h->host_ms, h->stage_ms[0], h->stage_ms[1], h->stage_ms[2], h->stage_ms[3], h->stage_ms[4]};
  const int m = n < 6 ? n : 6;
  for (int i = 0; i < m; i++) ms[i] = v[i];
  return m;
}

int aak_get_info(aak_handle h, int *info, int n) {
  if (!h) return fail(-1, "null handle");
  const int m = n < 10 ? n : 10;
  for (int i = 0; i < m; i++) info[i] = h->info[i];
  return m;
}

// ---------------------------------------------------------------- test hooks
int aak_test_host_plan(const double *params, double T_fit, int length, double dt, int LISA, int backint, int n_max,
                       double *traj_out, double *knots_out, int cap, int *info, double *dparams) {
  WaveConfig cfg;
  cfg.T_fit = T_fit;
  cfg.length = length;
  cfg.dt = dt;
  cfg.LISA = LISA;
  cfg.backint = backint;
  cfg.n_max = n_max;
  SourceParams sp;
  std::memcpy(&sp, params, sizeof(double) * 14);
  WaveformPlan plan;
  std::string err;
  const int rc = plan_waveform(cfg, sp, &plan, &err);
  if (rc < 0) return fail(rc, err);
  const SparseTrajectory &tr = plan.traj;
  for (int i = 0; i < 100; i++) {
    const bool ok = i < tr.steps;
    traj_out[i] = ok ? tr.e_traj[i] : 0;
    traj_out[100 + i] = ok ? tr.v_map[i] : 0;
    traj_out[200 + i] = ok ? tr.M_map[i] : 0;
    traj_out[300 + i] = ok ? tr.s_map[i] : 0;
    traj_out[400 + i] = ok ? tr.dt_map[i] : 0;
  }
  const KnotSet &ks = plan.knots;
  if (ks.K > cap) return fail(-6, "knot buffer too small");
  for (int k = 0; k < ks.K; k++) {
    knots_out[k] = ks.t[k];
    knots_out[cap + k] = ks.e[k];
    knots_out[2 * (size_t)cap + k] = ks.v[k];
    knots_out[3 * (size_t)cap + k] = ks.M[k];
    knots_out[4 * (size_t)cap + k] = ks.S[k];
  }
  int iters = 0;
  for (int i = 0; i < tr.steps; i++) iters += tr.map_iters[i];
  const int vals[16] = {tr.steps, ks.K, ks.i0, ks.j0, ks.j_min, ks.j_max, ks.j_max_temp, ks.j_start,
                        ks.j_end, ks.plunged, ks.points, plan.P.ip, plan.P.ib, plan.P.nmodes, plan.P.last_fold, iters};
  for (int i = 0; i < 16; i++) info[i] = vals[i];
  dparams[0] = tr.Phi0;
  dparams[1] = plan.P.zeta;
  dparams[2] = plan.P.t_plunge;
  dparams[3] = plan.P.t_zero;
  return ks.K;
}

int aak_test_spline_table(const double *params, double T_fit, int length, double dt, int LISA, int backint,
                          double *packed, int cap17) {
  WaveConfig cfg;
  cfg.T_fit = T_fit;
  cfg.length = length;
  cfg.dt = dt;
  cfg.LISA = LISA;
  cfg.backint = backint;
  SourceParams sp;
  std::memcpy(&sp, params, sizeof(double) * 14);
  WaveformPlan plan;
  std::string err;
  const int rc = plan_waveform(cfg, sp, &plan, &err);
  if (rc < 0) return fail(rc, err);
  if ((int)plan.packed.size() > cap17) return fail(-6, "table buffer too small");
  std::memcpy(packed, plan.packed.data(), sizeof(double) * plan.packed.size());
  return plan.knots.K;
}

int aak_test_inspiral_rates(double p, double e, double cosi, double a, double *out) {
  return inspiral_rates<HostCx>(p, e, cosi, a, out);
}

int aak_test_get_dense(aak_handle h, double *out) {
  if (!h) return fail(-1, "null handle");
  if (h->state == 0) return fail(-2, "no waveform generated");
  const size_t L = (size_t)h->N + 1;
  HIPCHK(hipMemcpyAsync(out, h->d_dense, sizeof(double) * 10 * L, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(hipStreamSynchronize(h->stream));
  return 0;
}

}  // extern "C"
