// aak_capi.cpp -- the C ABI (include/gpuAAK.h, include/aak_testing.h).
//
// One handle owns:
//   * a copy stream and one or more lanes (default priority streams), each
//     with the device buffers of one waveform in flight (dense arrays, phases,
//     h, t, fold and guard work, mode-sum coefficients, spectrum) and two
//     device knot tables, so the upload of the next waveform's table overlaps
//     the current one's kernels; lane 0 serves the single-waveform calls,
//     batches spread their templates over AAK_LANES lanes (default 2) so that
//     one template's narrow kernels overlap another's wide ones,
//   * the inner-product stage: for power-of-two lengths the fused four-step
//     FFT + weighted reduction (aak_fft.hip, data stored in its row order),
//     otherwise a rocFFT real-forward plan per lane (both channels in one
//     batched call) and the separate reduction; the Bessel tables,
//   * for batches, a pool of host threads and a ring of plan slots whose knot
//     tables live in pinned host memory.
//
// Per waveform a lane's stream runs, with no host round trip:
//   clear counters -> (wait upload) rates -> Phidot guard -> exact folds ->
//   fill -> mode sum [-> FFT + weighted reduction].
// gpu_gen_AAK / Likelihood keep the reference's one-waveform-per-call
// semantics; aak_likelihood_batch computes many likelihoods against the same
// data with the host stages of later templates running on the pool while the
// GPU works on earlier ones.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>
#include <sched.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/aak_testing.h"
#include "../../include/gpuAAK.h"
#include "aak_bessel.hpp"
#include "aak_dense.hpp"
#include "aak_fft.hpp"
#include "aak_freqmap.hpp"
#include "aak_geodesic.hpp"
#include "aak_host.hpp"
#include "aak_inspiral.hpp"
#include "aak_mlp.hpp"

using namespace aak;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(x)                                                                                    \
  do {                                                                                               \
    hipError_t e_ = (x);                                                                             \
    if (e_ != hipSuccess) return fail(-100, std::string("HIP: ") + hipGetErrorString(e_) + " at " #x); \
  } while (0)

std::once_flag g_fft_once;

// The host stage grows its evolution arrays with the knot count (2 j0
// doubles each, j0 up to 4e8): an allocation failure becomes a status code
// instead of an exception crossing the C boundary or ending a pool thread.
int plan_safe(const WaveConfig &cfg, const SourceParams &sp, WaveformPlan *plan, std::string *err,
              TableSink *sink = nullptr) noexcept {
  try {
    return plan_waveform(cfg, sp, plan, err, sink);
  } catch (const std::bad_alloc &) {
    *err = "host memory exhausted by the knot arrays";
  } catch (...) {
    *err = "host stage failed";
  }
  return -9;
}

// Idle pool workers pre-grow their evolution arrays to the largest plan seen,
// so that later batches do not allocate, but never beyond this many entries
// (7 arrays x 2^21 doubles = 117 MB per worker); a longer plan grows the
// arrays of the worker that runs it only.
constexpr size_t kIdleGrowCap = size_t(1) << 21;

// Threads the process may run on: its CPU affinity, bounded by a cgroup v2
// CPU quota when one is set.
int cpu_share() {
  int n = 0;
  cpu_set_t set;
  if (sched_getaffinity(0, sizeof set, &set) == 0) n = CPU_COUNT(&set);
  if (n <= 0) n = (int)std::thread::hardware_concurrency();
  if (FILE *f = std::fopen("/sys/fs/cgroup/cpu.max", "r")) {
    char q[32] = {0};
    long long period = 0;
    if (std::fscanf(f, "%31s %lld", q, &period) == 2 && std::strcmp(q, "max") != 0 && period > 0) {
      const long long quota = std::atoll(q);
      const int lim = (int)((quota + period - 1) / period);
      if (lim > 0 && lim < n) n = lim;
    }
    std::fclose(f);
  }
  return n > 0 ? n : 4;
}

// Pinned host buffer that grows on demand.
struct PinnedBuf {
  double *p = nullptr;
  size_t cap = 0;
  int reserve(size_t n) {
    if (n <= cap) return 0;
    release();
    const size_t c = n + n / 4 + 4096;
    if (hipHostMalloc((void **)&p, sizeof(double) * c, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return -1;
    }
    cap = c;
    return 0;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

struct PinnedSink : TableSink {
  PinnedBuf *buf;
  std::atomic<size_t> *hint;  // largest table seen by the pool (nullptr: none)
  explicit PinnedSink(PinnedBuf *b, std::atomic<size_t> *h = nullptr) : buf(b), hint(h) {}
  double *reserve(size_t n) override {
    if (n > buf->cap && hint) {
      // a slot that has to grow goes straight to the largest size seen, so each
      // slot of the ring allocates pinned memory at most about once
      size_t seen = hint->load();
      while (n > seen && !hint->compare_exchange_weak(seen, n)) {
      }
      if (seen > n) n = seen;
    }
    return buf->reserve(n) ? nullptr : buf->p;
  }
};

// One plan slot of the batch pool.  Workers take a free slot, fill it with
// the plan of the next template and queue it; the consumer takes queued
// slots in completion order (a slow host stage does not hold up the GPU) and
// frees them once their table upload is enqueued.
struct RingSlot {
  DenseParams P;                // plan of the template held
  size_t tab_n = 0;             // knot table length (in pin)
  PinnedBuf pin;
  hipEvent_t copied = nullptr;  // upload of this slot's table finished
  bool copy_pending = false;
  long long tpl = -1;           // template index held
  int rc = 0;
  double plan_ms = 0.0;         // host stage time of the template held
  ~RingSlot() {
    pin.release();
    if (copied) (void)hipEventDestroy(copied);
  }
};

}  // namespace

// Device state of one waveform in flight: a compute stream with its own dense
// arrays, fold work, mode-sum coefficients, time-domain and spectrum buffers,
// two knot-table slots and a rocFFT execution context.  Batches spread their
// templates over several lanes so that one template's serial or narrow
// kernels (the fold chain, small reductions) overlap another's wide ones.
struct Lane {
  hipStream_t stream = nullptr;
  double *d_table[2] = {nullptr, nullptr};
  size_t table_cap[2] = {0, 0};
  hipEvent_t ev_upload[2] = {nullptr, nullptr}, ev_rates[2] = {nullptr, nullptr};
  bool rates_recorded[2] = {false, false};
  double *d_dense = nullptr;  // e, v, M, S, rate[3], phase[3], amp: 11 x (N+1), then nu0
  double *d_h = nullptr;      // hI | hII, N+2 samples each at channel stride h_stride(N)
  double *d_t = nullptr;      // [N+2]
  int *d_guardw = nullptr;    // [kGuardGrid + 1] (DenseBuffers::guard_work)
  int *d_flags = nullptr;     // [8]: non-positive Phidot count, fold fallbacks, ... (DenseBuffers::flags)
  bool flags_dirty = false;   // a waveform wrote d_flags and no final reduction has zeroed them since
  FoldWork fw{};       // forward fold (and both directions of the three-pass form)
  FoldWork fw_back{};  // backward direction of the single-pass fold
  FoldReset fold_reset{};
  int fold_nblk_cap = 0;
  void *d_foldmem = nullptr;
  double2 *d_hf = nullptr;    // [2][nbins]
  double2 *d_fy = nullptr;    // [2][N/2] four-step FFT scratch (fused FFT path)
  double *d_partials = nullptr, *d_like = nullptr;
  double *d_fparts = nullptr;     // pass 2 with the final reduction: published pairs
  double *d_ms_coef = nullptr;  // mode-sum block Taylor coefficients (grown with nmodes)
  size_t ms_coef_cap = 0;
  int *d_ms_p = nullptr;
  int *d_ms_slot = nullptr;
  int *d_ms_list = nullptr;
  int *d_ms_split = nullptr;
  double *d_ms_eref = nullptr;
  rocfft_plan plan = nullptr;
  rocfft_execution_info exec = nullptr;
  void *d_work = nullptr;
  hipGraph_t g_like = nullptr;          // the batch's likelihood stage (fused FFT), captured on first use
  hipGraphExec_t gx_like = nullptr;
  bool ready = false;
};

constexpr int kMaxLanes = 8;
// short plans kept ready or in progress for the GPU by the batch pool (about
// 3 ms of device work at N = 2^21); other picks take the longest remaining.
// Measured over 256-template batches of the production prior on 16 host
// threads: 6 -> 3.26e3, 8 -> 3.42e3, 10 -> 3.41e3 (FP64 +3%), 12 -> 3.09e3,
// 16 -> 2.63e3 likelihoods/s (too few threads left for the longest plans).
constexpr int kFeedDepth = 10;  // default of aak_handle_s::feed_depth

struct aak_handle_s {
  WaveConfig cfg;
  int N = 0, nbins = 0;
  hipStream_t cstream = nullptr;  // knot-table uploads of every lane
  Lane lane[kMaxLanes];
  int nlanes = 2;                 // lanes used by batches (lane 0 serves single calls)
  int feed_depth = kFeedDepth;    // short plans kept ready or in progress for the GPU
  bool use_graphs = true;         // batches replay each lane's captured likelihood stage (AAK_GRAPHS=0: off)
  bool fuse_final = true;         // pass 2 ends in the final reduction (AAK_FUSED_FINAL=0: separate launch)
  double *d_bessel = nullptr;     // c_m(k) table | log k!
  double2 *d_data = nullptr;      // [2][nbins] weighted data (fused FFT path: in its row order)
  double *d_w = nullptr;          // [2][nbins] weights (same order as d_data)
  bool fused_fft = false;         // power-of-two length: fused FFT + reduction, else rocFFT
  FftArgs fft{};
  double2 *d_fft_tw = nullptr;    // twiddle tables of the fused FFT
  double *d_like_batch = nullptr;
  double *d_phase = nullptr;      // [8][N] phase output, allocated on first use
  size_t like_batch_cap = 0;
  PinnedBuf like_host;
  int have_data = 0;
  int state = 0;  // 0 empty, 1 time domain, 2 weighted spectrum (lane 0)
  int info[12] = {0};
  // single-waveform staging
  WaveformPlan single;
  PinnedBuf single_pin;
  int have_traj = 0;  // single holds a whole host stage (run_phase_trajectory, gpu_gen_AAK)
  // M (solar masses), cos(iota) and plunge index of the last gpu_gen_AAK,
  // whose dense arrays lane 0 keeps while state != 0
  double gen_Mphys = 0.0, gen_coslam = 1.0;
  int gen_ip = -1;
  // stage timing on the compute stream: rates, guard+folds, mode sum, rFFT, reduction
  int profiling = 0;
  int force_fold_replay = 0;  // test hook: every fold takes the serial replay
  int fold_mode = 0;          // test hook: 0 single-pass look-back fold, 1 three-pass fold
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  hipEvent_t *evs = ev;               // events the marks go to (per template in a batch)
  std::vector<hipEvent_t> batch_ev;   // 6 per template of the last profiled batch
  double stage_ms[5] = {0, 0, 0, 0, 0};
  double batch_ms[5] = {0, 0, 0, 0, 0};  // stage sums over the last profiled batch
  double batch_host_sum = 0.0, batch_host_max = 0.0;  // host stage ms over the last batch
  double batch_wall_ms = 0.0;                         // wall time of the last batch call
  double batch_issue_ms = 0.0, batch_wait_ms = 0.0;   // issuing thread: enqueueing, waiting for plans
  int batch_profiled = 0;
  double host_ms = 0.0;
  // batch machinery
  int host_threads = 0;
  std::vector<std::thread> workers;
  std::vector<std::unique_ptr<RingSlot>> ring;
  std::vector<int> free_slots;
  std::deque<int> ready_slots;
  std::mutex mu;
  std::condition_variable cv_work, cv_ready;
  bool stop = false;
  const double *b_params = nullptr;
  const int *b_nmax = nullptr;                // per-template harmonic counts (optional)
  long long b_n = 0, b_next = 0, b_end = 0;  // b_order[b_next, b_end) not yet taken
  std::vector<long long> b_order;  // templates in decreasing estimated host cost
  int cheap_inflight = 0;          // plans in progress taken from the cheap end (under mu)
  std::atomic<size_t> pin_hint{0};  // largest knot table (doubles) planned by the pool
  size_t knot_hint = 0;             // largest evolution array length planned by the pool (under mu)
  std::vector<size_t> worker_cap;   // evolution array length of each worker (under mu)

  void mark(const Lane &L, int k) {
    if (profiling) (void)hipEventRecord(evs[k], L.stream);
  }
  DenseBuffers bufs(const Lane &ln, int slot) const {
    DenseBuffers B;
    const size_t L = (size_t)N + 1;
    B.knots = ln.d_table[slot];
    B.e = ln.d_dense;
    B.v = ln.d_dense + L;
    B.M = ln.d_dense + 2 * L;
    B.S = ln.d_dense + 3 * L;
    for (int p = 0; p < 3; p++) {
      B.rate[p] = ln.d_dense + (4 + p) * L;
      B.phase[p] = ln.d_dense + (7 + p) * L;
    }
    B.amp = ln.d_dense + 10 * L;
    B.nu0 = ln.d_dense + 11 * L;
    B.h = ln.d_h;
    B.h2 = ln.d_h + h_stride(N);
    B.t = ln.d_t;
    B.flags = ln.d_flags;
    B.guard_work = ln.d_guardw;
    B.ms_coef = ln.d_ms_coef;
    B.ms_p = ln.d_ms_p;
    B.ms_slot = ln.d_ms_slot;
    B.ms_list = ln.d_ms_list;
    B.ms_split = ln.d_ms_split;
    B.ms_eref = ln.d_ms_eref;
    B.bessel_c = d_bessel;
    B.bessel_lgf = d_bessel + (size_t)(kBesselKMax + 1) * kBesselMMax;
    return B;
  }
};

static int alloc_lane(aak_handle h, Lane &L) {
  if (L.ready) return 0;
  const int N = h->N;
  HIPCHK(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  for (int k = 0; k < 2; k++) {
    HIPCHK(hipEventCreateWithFlags(&L.ev_upload[k], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&L.ev_rates[k], hipEventDisableTiming));
  }
  HIPCHK(hipMalloc(&L.d_dense, sizeof(double) * (11 * ((size_t)N + 1) + 1)));
  HIPCHK(hipMalloc(&L.d_h, sizeof(double) * 2 * h_stride(N)));
  HIPCHK(hipMalloc(&L.d_t, sizeof(double) * ((size_t)N + 2)));
  HIPCHK(hipMalloc(&L.d_flags, sizeof(int) * 8));
  HIPCHK(hipMalloc(&L.d_guardw, sizeof(int) * (kGuardGrid + 1)));
  HIPCHK(hipMemset(L.d_flags, 0, sizeof(int) * 8));
  HIPCHK(hipMalloc(&L.d_ms_p, sizeof(int) * ((size_t)N / 256 + 4)));
  HIPCHK(hipMalloc(&L.d_ms_slot, sizeof(int) * ((size_t)N / 256 + 4)));
  HIPCHK(hipMalloc(&L.d_ms_list, sizeof(int) * ((size_t)N / 256 + 4)));
  HIPCHK(hipMalloc(&L.d_ms_split, sizeof(int) * ((size_t)N / 256 + 4)));
  HIPCHK(hipMalloc(&L.d_ms_eref, sizeof(double) * 2 * ((size_t)N / 256 + 4)));
  const int nblk = (N + kFoldBlock - 1) / kFoldBlock + 1;
  L.fold_nblk_cap = nblk;
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
  const size_t o_st = carve(sizeof(int) * 6);
  const size_t o_lw = carve(sizeof(long long) * 2 * 3 * (size_t)nblk * kLW);
  const size_t o_lc = carve(sizeof(unsigned) * 4);
  HIPCHK(hipMalloc(&L.d_foldmem, bytes));
  char *base = (char *)L.d_foldmem;
  L.fw.bsum = (double *)(base + o_bsum);
  L.fw.bmin = (double *)(base + o_bmin);
  L.fw.bmax = (double *)(base + o_bmax);
  L.fw.xstart = (double *)(base + o_xs);
  L.fw.E = (double *)(base + o_E);
  L.fw.nitems = (int *)(base + o_ni);
  L.fw.nties = (int *)(base + o_nt);
  L.fw.items = (FoldItem *)(base + o_items);
  L.fw.ties = (FoldTie *)(base + o_ties);
  L.fw.status = (int *)(base + o_st);
  L.fw.look = (long long *)(base + o_lw);
  L.fw.look_ctr = (unsigned *)(base + o_lc);
  L.fw.fallbacks = L.d_flags + 1;
  // the backward direction of the single-pass fold has words of its own
  L.fw_back = L.fw;
  L.fw_back.status = L.fw.status + 3;
  L.fw_back.look = L.fw.look + 3 * (size_t)nblk * kLW;
  L.fw_back.look_ctr = L.fw.look_ctr + 1;
  L.fold_reset = FoldReset{L.fw.look, (int)(2 * 3 * (size_t)nblk * kLW), L.fw.status, L.fw.look_ctr, 0};
  const int nb = h->nbins;
  HIPCHK(hipMalloc(&L.d_hf, sizeof(double2) * 2 * (size_t)nb));
  const int nparts = std::max(2048, h->fused_fft ? fft_partials(h->fft) : 0);
  HIPCHK(hipMalloc(&L.d_partials, sizeof(double) * 2 * (size_t)nparts));
  HIPCHK(hipMalloc(&L.d_like, sizeof(double) * 2));
  if (h->fused_fft) {
    HIPCHK(hipMalloc(&L.d_fy, sizeof(double2) * 2 * (size_t)h->fft.M));
    const int fp = fft_partials(h->fft);
    HIPCHK(hipMalloc(&L.d_fparts, sizeof(double) * 2 * (size_t)fp));
    launch_fft_final_reset(L.d_fparts, fp, L.stream);
    HIPCHK(hipGetLastError());
    L.ready = true;
    return 0;
  }
  // rocFFT: two real sequences of length N (distance N+2) -> two half spectra
  std::call_once(g_fft_once, []() { rocfft_setup(); });
  rocfft_plan_description desc = nullptr;
  if (rocfft_plan_description_create(&desc) != rocfft_status_success) return fail(-101, "rocfft description");
  const size_t in_stride = 1, out_stride = 1;
  const size_t in_dist = h_stride(N), out_dist = (size_t)nb;
  if (rocfft_plan_description_set_data_layout(desc, rocfft_array_type_real, rocfft_array_type_hermitian_interleaved,
                                              nullptr, nullptr, 1, &in_stride, in_dist, 1, &out_stride,
                                              out_dist) != rocfft_status_success) {
    rocfft_plan_description_destroy(desc);
    return fail(-101, "rocfft layout");
  }
  const size_t len = (size_t)N;
  const rocfft_status ps = rocfft_plan_create(&L.plan, rocfft_placement_notinplace, rocfft_transform_type_real_forward,
                                              rocfft_precision_double, 1, &len, 2, desc);
  rocfft_plan_description_destroy(desc);
  if (ps != rocfft_status_success) return fail(-101, "rocfft plan");
  size_t wsize = 0;
  rocfft_plan_get_work_buffer_size(L.plan, &wsize);
  if (rocfft_execution_info_create(&L.exec) != rocfft_status_success) return fail(-101, "rocfft execution info");
  if (wsize) {
    HIPCHK(hipMalloc(&L.d_work, wsize));
    rocfft_execution_info_set_work_buffer(L.exec, L.d_work, wsize);
  }
  rocfft_execution_info_set_stream(L.exec, L.stream);
  L.ready = true;
  return 0;
}

static void free_lane(Lane &L) {
  if (L.stream) (void)hipStreamSynchronize(L.stream);
  if (L.gx_like) (void)hipGraphExecDestroy(L.gx_like);
  if (L.g_like) (void)hipGraphDestroy(L.g_like);
  if (L.plan) rocfft_plan_destroy(L.plan);
  if (L.exec) rocfft_execution_info_destroy(L.exec);
  void *ptrs[] = {L.d_table[0], L.d_table[1], L.d_dense, L.d_h, L.d_t, L.d_flags, L.d_guardw, L.d_foldmem,
                  L.d_hf, L.d_fy, L.d_partials, L.d_fparts, L.d_like, L.d_work, L.d_ms_coef, L.d_ms_p, L.d_ms_slot,
                  L.d_ms_list, L.d_ms_split, L.d_ms_eref};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (int k = 0; k < 2; k++) {
    if (L.ev_upload[k]) (void)hipEventDestroy(L.ev_upload[k]);
    if (L.ev_rates[k]) (void)hipEventDestroy(L.ev_rates[k]);
  }
  if (L.stream) (void)hipStreamDestroy(L.stream);
  L = Lane();
}

static int alloc_handle(aak_handle h) {
  HIPCHK(hipStreamCreateWithFlags(&h->cstream, hipStreamNonBlocking));
  {
    const size_t nc = (size_t)(kBesselKMax + 1) * kBesselMMax, nl = kBesselKMax + 1;
    std::vector<double> tab(nc + nl);
    bessel_series_table(tab.data());
    bessel_log_factorials(tab.data() + nc);
    HIPCHK(hipMalloc(&h->d_bessel, sizeof(double) * tab.size()));
    HIPCHK(hipMemcpy(h->d_bessel, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
  }
  const int nb = h->nbins;
  HIPCHK(hipMalloc(&h->d_data, sizeof(double2) * 2 * (size_t)nb));
  HIPCHK(hipMalloc(&h->d_w, sizeof(double) * 2 * (size_t)nb));
  if (h->fused_fft) {
    FftArgs &F = h->fft;
    std::vector<double2> tw(fft_table_size(F));
    fft_fill_tables(F, tw.data());
    HIPCHK(hipMalloc(&h->d_fft_tw, sizeof(double2) * tw.size()));
    HIPCHK(hipMemcpy(h->d_fft_tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice));
    fft_bind_tables(&F, h->d_fft_tw);
    F.Dp[0] = h->d_data;
    F.Dp[1] = h->d_data + nb;
    F.wp[0] = h->d_w;
    F.wp[1] = h->d_w + nb;
  }
  return alloc_lane(h, h->lane[0]);
}

static void stop_workers(aak_handle h) {
  {
    std::lock_guard<std::mutex> lk(h->mu);
    h->stop = true;
  }
  h->cv_work.notify_all();
  h->cv_ready.notify_all();
  for (std::thread &t : h->workers)
    if (t.joinable()) t.join();
  h->workers.clear();
  h->stop = false;
  if (h->cstream) (void)hipStreamSynchronize(h->cstream);
  h->ring.clear();
}

static void free_handle(aak_handle h) {
  stop_workers(h);
  if (h->cstream) (void)hipStreamSynchronize(h->cstream);
  for (Lane &L : h->lane) free_lane(L);
  void *ptrs[] = {h->d_bessel, h->d_data, h->d_w, h->d_fft_tw, h->d_like_batch, h->d_phase};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  h->single_pin.release();
  h->like_host.release();
  for (hipEvent_t &e : h->ev)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t &e : h->batch_ev)
    if (e) (void)hipEventDestroy(e);
  if (h->cstream) (void)hipStreamDestroy(h->cstream);
}

// Enqueue the device stage of one planned waveform on lane L.  The knot table
// goes to the lane's table `slot` on the copy stream; `copied` (optional) is
// recorded when the host table may be reused.
static int enqueue_waveform(aak_handle h, Lane &L, const DenseParams &Pplan, const double *tab, size_t tab_n, int slot,
                            hipEvent_t copied, const DenseParams *Pover = nullptr) {
  const DenseParams &P = Pover ? *Pover : Pplan;
  const size_t need = tab_n;
  if (need > L.table_cap[slot]) {
    // growing: no kernel may still read the old table
    HIPCHK(hipStreamSynchronize(L.stream));
    HIPCHK(hipStreamSynchronize(h->cstream));
    if (L.d_table[slot]) HIPCHK(hipFree(L.d_table[slot]));
    L.d_table[slot] = nullptr;
    L.table_cap[slot] = 0;
    const size_t cap = need + need / 4 + 4096;
    HIPCHK(hipMalloc(&L.d_table[slot], sizeof(double) * cap));
    L.table_cap[slot] = cap;
    L.rates_recorded[slot] = false;
  }
  if (!Pover) {
    // block Taylor coefficients: one slot per expansion block and one per
    // 256-sample block (split expansions), nmodes x 3 x (kTaylorMax+1) doubles each
    const size_t nc = (size_t)modesum_slots(P) * (size_t)(P.nmodes > 0 ? P.nmodes : 1) * 3 * (kTaylorMax + 1);
    if (nc > L.ms_coef_cap) {
      HIPCHK(hipStreamSynchronize(L.stream));
      if (L.d_ms_coef) HIPCHK(hipFree(L.d_ms_coef));
      L.d_ms_coef = nullptr;
      L.ms_coef_cap = 0;
      HIPCHK(hipMalloc(&L.d_ms_coef, sizeof(double) * nc));
      L.ms_coef_cap = nc;
    }
  }
  if (L.rates_recorded[slot]) HIPCHK(hipStreamWaitEvent(h->cstream, L.ev_rates[slot], 0));
  HIPCHK(hipMemcpyAsync(L.d_table[slot], tab, sizeof(double) * need, hipMemcpyHostToDevice, h->cstream));
  HIPCHK(hipEventRecord(L.ev_upload[slot], h->cstream));
  if (copied) HIPCHK(hipEventRecord(copied, h->cstream));
  const DenseBuffers B = h->bufs(L, slot);
  if (L.flags_dirty) launch_clear_flags(B, L.stream);
  L.flags_dirty = true;  // the waveform's kernels write them
  HIPCHK(hipStreamWaitEvent(L.stream, L.ev_upload[slot], 0));
  h->mark(L, 0);
  launch_rates(P, B, L.stream);
  HIPCHK(hipEventRecord(L.ev_rates[slot], L.stream));
  L.rates_recorded[slot] = true;
  h->mark(L, 1);
  {
    FoldReset R = L.fold_reset;
    R.force = h->force_fold_replay;
    launch_phidot_guard(P, B, R, L.stream);
  }
  const int nfwd = P.last_fold - P.i0 + 1;
  FoldWork fws[2] = {L.fw, h->fold_mode == 0 ? L.fw_back : L.fw};
  fws[0].force_replay = fws[1].force_replay = h->force_fold_replay;
  // the replay check of the last fold direction rides on the mode-sum setup
  // when one follows (the two directions share the fold's work arrays, so an
  // earlier direction keeps its own launch)
  FoldArgs deferred{};
  bool have_deferred = false;
  int deferred_pass = 0;
  for (int pass = 0; pass < 2; pass++) {
    const int dir = pass == 0 ? 1 : -1;
    const int nsteps = pass == 0 ? nfwd : P.i0;
    if (nsteps <= 0) continue;
    const bool last = pass == 1 || P.i0 <= 0;
    const bool defer = last && !Pover && h->fold_mode == 0;
    if (fold_exact(P, B, fws[pass], L.fold_nblk_cap, dir, nsteps, h->fold_mode, defer ? &deferred : nullptr,
                   L.stream))
      return fail(-7, "fold block capacity");
    have_deferred = defer;
    deferred_pass = pass;
  }
  launch_fold_fill(P, B, L.stream);
  h->mark(L, 2);
  if (Pover) {  // dense stage only
    HIPCHK(hipGetLastError());
    return 0;
  }
  launch_modesum(P, B, have_deferred ? &deferred : nullptr, &fws[deferred_pass], L.stream);
  h->mark(L, 3);
  HIPCHK(hipGetLastError());
  return 0;
}

// flags: the lane's waveform flags when the series is a generated waveform (a
// waveform rejected on the device yields NaN), nullptr for a given series.
// The batch's two FFT passes captured once per lane as a HIP graph: every
// pointer they use belongs to the lane or the handle, so one graph serves
// every template of the lane.  It holds kernels only (no memset nodes).  The
// final reduction is launched after it with the template's own output row.
static int lane_like_graph(aak_handle h, Lane &L) {
  if (L.gx_like) return 0;
  HIPCHK(hipStreamBeginCapture(L.stream, hipStreamCaptureModeThreadLocal));
  launch_fft_pass1(h->fft, L.d_h, L.d_h + h_stride(h->N), L.d_fy, L.stream);
  launch_fft_pass2(h->fft, L.d_fy, L.d_partials, nullptr, L.stream);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(L.stream, &g);
  if (e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(-100, std::string("HIP: ") + hipGetErrorString(e) + " capturing the likelihood stage");
  }
  L.g_like = g;
  HIPCHK(hipGraphInstantiate(&L.gx_like, g, nullptr, nullptr, 0));
  return 0;
}

// The final reduction zeroes the lane's flags once read when they are the
// flags of the waveform in hand, so the lane's next waveform needs no
// clearing launch (Lane::flags_dirty).
static int enqueue_likelihood(aak_handle h, Lane &L, double *out_dev, bool keep_spectrum, const int *flags) {
  int *clear = (flags == L.d_flags) ? L.d_flags : nullptr;
  if (h->fused_fft && h->fuse_final && !keep_spectrum) {
    // two launches: the second pass's last workgroup writes the template's row
    launch_fft_pass1(h->fft, L.d_h, L.d_h + h_stride(h->N), L.d_fy, L.stream);
    h->mark(L, 4);
    launch_fft_pass2_final(h->fft, L.d_fy, L.d_fparts, FftFinal{out_dev, flags, clear}, L.stream);
    h->mark(L, 5);
    HIPCHK(hipGetLastError());
    if (clear) L.flags_dirty = false;
    return 0;
  }
  if (h->fused_fft && h->use_graphs && !keep_spectrum && !h->profiling && clear) {
    int rc = lane_like_graph(h, L);
    if (rc) return rc;
    HIPCHK(hipGraphLaunch(L.gx_like, L.stream));
    launch_dot_final(L.d_partials, fft_partials(h->fft), out_dev, flags, clear, L.stream);
    HIPCHK(hipGetLastError());
    L.flags_dirty = false;
    return 0;
  }
  if (h->fused_fft) {
    launch_fft_pass1(h->fft, L.d_h, L.d_h + h_stride(h->N), L.d_fy, L.stream);
    h->mark(L, 4);
    launch_fft_pass2(h->fft, L.d_fy, L.d_partials, keep_spectrum ? L.d_hf : nullptr, L.stream);
    launch_dot_final(L.d_partials, fft_partials(h->fft), out_dev, flags, clear, L.stream);
    h->mark(L, 5);
    HIPCHK(hipGetLastError());
    if (clear) L.flags_dirty = false;
    return 0;
  }
  void *in[1] = {L.d_h};
  void *outb[1] = {L.d_hf};
  if (rocfft_execute(L.plan, in, outb, L.exec) != rocfft_status_success) return fail(-101, "rocfft execute");
  h->mark(L, 4);
  launch_weighted_dot(L.d_hf, h->d_data, h->d_w, h->nbins, keep_spectrum ? L.d_hf : nullptr, L.d_partials, out_dev,
                      flags, clear, L.stream);
  h->mark(L, 5);
  HIPCHK(hipGetLastError());
  if (clear) L.flags_dirty = false;
  return 0;
}

// stage k spans events k .. k+1
static void read_stage_times(aak_handle h, int first, int last) {
  if (!h->profiling) return;
  float ms = 0.f;
  for (int k = first; k < last; k++)
    if (hipEventElapsedTime(&ms, h->ev[k], h->ev[k + 1]) == hipSuccess) h->stage_ms[k] = ms;
}

// ---------------------------------------------------------------- batch pool
static void worker_main(aak_handle h, int id) {
  // the evolution arrays of the worker's plans, reused across templates and
  // batches; the slot keeps the parameters and the knot table
  WaveformPlan plan;
  size_t grow_failed = SIZE_MAX;  // idle growth to this size failed: not retried
  for (;;) {
    long long i;
    const double *params;
    const int *nmax_rows = nullptr;
    RingSlot *slot;
    int sidx;
    size_t grow_to = 0;
    bool took_cheap = false;
    {
      std::unique_lock<std::mutex> lk(h->mu);
      auto has_work = [&] { return h->b_params && h->b_next < h->b_end && !h->free_slots.empty(); };
      auto idle_target = [&] { return std::min(h->knot_hint, kIdleGrowCap); };
      auto wants_growth = [&] { return plan.knots.eA.size() < idle_target() && idle_target() < grow_failed; };
      h->cv_work.wait(lk, [&] { return h->stop || has_work() || wants_growth(); });
      if (h->stop) return;
      if (has_work()) {
        // keep about kFeedDepth short plans ready or in progress for the GPU;
        // otherwise work down from the longest, which bound the batch
        took_cheap = (int)h->ready_slots.size() + h->cheap_inflight < h->feed_depth;
        if (took_cheap) h->cheap_inflight++;
        i = took_cheap ? h->b_order[(size_t)--h->b_end] : h->b_order[(size_t)h->b_next++];
        params = h->b_params;
        nmax_rows = h->b_nmax;
        sidx = h->free_slots.back();
        h->free_slots.pop_back();
        slot = h->ring[(size_t)sidx].get();
      } else {
        grow_to = idle_target();
      }
    }
    if (grow_to) {
      // idle: grow to the largest evolution planned so far (up to
      // kIdleGrowCap), outside any template; on failure the worker's next plan
      // grows what it needs, or reports the failure
      bool grown = true;
      try {
        reserve_knots(&plan.knots, grow_to);
      } catch (const std::bad_alloc &) {
        grown = false;
        grow_failed = grow_to;
      }
      {
        std::lock_guard<std::mutex> lk(h->mu);
        // a failed growth reports the target as reached, so the end of a batch
        // does not wait for it
        h->worker_cap[(size_t)id] = grown ? plan.knots.eA.size() : grow_to;
      }
      h->cv_ready.notify_all();
      continue;
    }
    if (slot->copy_pending) {
      (void)hipEventSynchronize(slot->copied);
      slot->copy_pending = false;
    }
    SourceParams sp;
    std::memcpy(&sp, params + 14 * (size_t)i, sizeof(SourceParams));
    PinnedSink sink(&slot->pin, &h->pin_hint);
    std::string err;
    const auto t_plan = std::chrono::steady_clock::now();
    WaveConfig cfg = h->cfg;
    if (nmax_rows) cfg.n_max = nmax_rows[i];
    const int rc = plan_safe(cfg, sp, &plan, &err, &sink);
    if (rc == 0) {
      // likelihoods need the spline only up to the end of the signal
      plan.P.full_dense = 0;
      trim_table_to_signal(&plan);
    }
    slot->plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_plan).count();
    slot->P = plan.P;
    slot->tab_n = plan.tab_n;
    bool hint_grew = false;
    {
      std::lock_guard<std::mutex> lk(h->mu);
      h->worker_cap[(size_t)id] = plan.knots.eA.size();
      if (plan.knots.eA.size() > h->knot_hint) {
        h->knot_hint = plan.knots.eA.size();
        hint_grew = true;
      }
      slot->rc = rc;
      slot->tpl = i;
      h->ready_slots.push_back(sidx);
      if (took_cheap) h->cheap_inflight--;
    }
    h->cv_ready.notify_all();
    if (hint_grew) h->cv_work.notify_all();  // idle workers grow to the new size
  }
}

static int start_pool(aak_handle h) {
  if (!h->workers.empty()) return 0;
  int T = h->host_threads;
  if (T <= 0) {
    const char *env = std::getenv("AAK_HOST_THREADS");
    T = env ? std::atoi(env) : 0;
    if (T <= 0) T = cpu_share();
  }
  const int R = T + 4;
  h->ring.clear();
  for (int s = 0; s < R; s++) {
    std::unique_ptr<RingSlot> slot(new RingSlot());
    HIPCHK(hipEventCreateWithFlags(&slot->copied, hipEventDisableTiming));
    h->ring.push_back(std::move(slot));
  }
  h->worker_cap.assign((size_t)T, 0);
  for (int t = 0; t < T; t++) h->workers.emplace_back(worker_main, h, t);
  return 0;
}

namespace aak {
// The calling thread's error message, for entry points in other files.
int set_last_error(int code, const std::string &msg) { return fail(code, msg); }
}  // namespace aak

extern "C" {

const char *aak_last_error(void) { return g_err.c_str(); }

double aak_OmegaPhi(double v, double e, double cosiota, double s, double M) { return omega_phi(v, e, cosiota, s, M); }

// Knot outputs of one host stage: the reference's PNevolution results
// (gpuAAK.cc:234-266) from a knot set and its table.
static int knot_outputs(const WaveConfig &cfg, const KnotSet &ks, const SparseTrajectory &tr, const double *par,
                        double *t_in, double *e_in, double *v_in, double *M_in, double *S_in, int cap, int *i_plunge,
                        int *i_buffer, double *t_clip, double *interp_timestep, int *init_length) {
  const int K = ks.K;
  if (K > cap) return fail(-6, "knot buffer too small (" + std::to_string(K) + " knots)");
  std::vector<double> tab(table_size(ks));
  KnotGrid g;
  fill_table(ks, tr, tab.data(), &g);
  int ip = 0, ib = 0;
  double tp = 0.0;
  std::string err;
  const int rc = plunge_buffer(cfg, ks, tab.data(), g, cos(par[4]), &ip, &ib, &tp, &err);
  if (rc < 0) return fail(rc, err);
  double *dst[5] = {t_in, e_in, v_in, M_in, S_in};
  std::memcpy(dst[0], ks.kt(), sizeof(double) * K);
  for (int a = 0; a < 4; a++) std::memcpy(dst[a + 1], ks.ky(a), sizeof(double) * K);
  *i_plunge = ip - ks.i0;
  *i_buffer = ib;
  *t_clip = (ip + ib - ks.i0) * cfg.dt;
  if (*t_clip > ks.kt()[K - 1]) *t_clip = ks.kt()[K - 1];
  *interp_timestep = ks.dt_large;
  *init_length = K;
  return 0;
}

int aak_PNevolution(double *t_in, double *e_in, double *v_in, double *M_in, double *S_in, int cap, double timestep,
                    int vlength, const double *par, const double *e_traj, const double *v_map, double M_phys,
                    const double *M_map, double S_phys, const double *S_map, const double *dt_map, int steps,
                    int *i_plunge, int *i_buffer, int backint, double *t_clip, double *interp_timestep,
                    int *init_length) {
  if (!t_in || !e_in || !v_in || !M_in || !S_in || !par || !e_traj || !v_map || !M_map || !S_map || !dt_map ||
      !i_plunge || !i_buffer || !t_clip || !interp_timestep || !init_length)
    return fail(-2, "null argument");
  if (steps < 3 || steps > 100) return fail(-2, "steps must lie in [3, 100]");
  if (vlength < 2 || !(timestep > 0) || cap < 1) return fail(-2, "bad grid or capacity");
  WaveConfig cfg;
  cfg.length = vlength;
  cfg.dt = timestep;
  cfg.backint = backint ? 1 : 0;
  SparseTrajectory tr;
  tr.steps = steps;
  std::memcpy(tr.e_traj, e_traj, sizeof(double) * steps);
  std::memcpy(tr.v_map, v_map, sizeof(double) * steps);
  std::memcpy(tr.M_map, M_map, sizeof(double) * steps);
  std::memcpy(tr.s_map, S_map, sizeof(double) * steps);
  std::memcpy(tr.dt_map, dt_map, sizeof(double) * steps);
  try {
    KnotSet ks;
    std::string err;
    const int rc = build_knots(cfg, par, M_phys, S_phys, tr, &ks, &err);
    if (rc < 0) return fail(rc, err);
    return knot_outputs(cfg, ks, tr, par, t_in, e_in, v_in, M_in, S_in, cap, i_plunge, i_buffer, t_clip,
                        interp_timestep, init_length);
  } catch (const std::bad_alloc &) {
    return fail(-9, "host memory exhausted by the knot arrays");
  }
}

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
  if (const char *env = std::getenv("AAK_GRAPHS")) h->use_graphs = std::atoi(env) != 0;
  if (const char *env = std::getenv("AAK_FUSED_FINAL")) h->fuse_final = std::atoi(env) != 0;
  if (const char *env = std::getenv("AAK_LANES")) {
    const int v = std::atoi(env);
    h->nlanes = v < 1 ? 1 : (v > kMaxLanes ? kMaxLanes : v);
  }
  {
    // AAK_FFT=rocfft keeps the library transform for every length
    const char *env = std::getenv("AAK_FFT");
    int R = 0, C = 0;
    if (!(env && std::strcmp(env, "rocfft") == 0) && fft_supported(length, &R, &C)) {
      h->fused_fft = true;
      h->fft.N = length;
      h->fft.M = length / 2;
      h->fft.R = R;
      h->fft.C = C;
    }
  }
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

int aak_set_map_init(aak_handle h, int mode, aak_mlp m) {
  if (!h) return fail(-1, "null handle");
  if (mode != 0 && mode != 1) return fail(-2, "map init mode must be 0 or 1");
  if (mode == 1 && !m) return fail(-2, "map init mode 1 needs a model");
  h->cfg.map_init = mode;
  h->cfg.mlp = mode == 1 ? m : nullptr;
  return 0;
}

double aak_p_lso(double a, double e, double iota) {
  Inspiral<HostCx> insp;
  if (insp.init(a)) return std::nan("");
  return insp.p_lso(cos(iota), e);
}

int aak_p_lso_batch(const double *aei, int n, double *out, int threads) {
  if (n < 0 || (n > 0 && (!aei || !out))) return fail(-2, "bad arguments");
  int T = threads > 0 ? threads : cpu_share();
  T = std::max(1, std::min(T, n));
  auto work = [&](int id) {
    for (long long i = id; i < n; i += T) out[i] = aak_p_lso(aei[3 * i], aei[3 * i + 1], aei[3 * i + 2]);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; t++) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  return 0;
}

int aak_freqmap_batch(const double *pts, int n, int mode, aak_mlp m, double tol, int threads, double *maps,
                      int *iters, double *resid, double *stats) {
  if (n < 0 || (n > 0 && (!pts || !maps))) return fail(-2, "bad arguments");
  if (mode != 0 && mode != 1) return fail(-2, "mode must be 0 or 1");
  if (mode == 1 && !m) return fail(-2, "mode 1 needs a model");
  if (!(tol > 0)) return fail(-2, "tolerance must be positive");
  if (stats)
    for (int k = 0; k < 4; k++) stats[k] = 0.0;
  if (n == 0) return 0;
  std::vector<double> starts;
  if (mode == 1) {
    starts.resize(3 * (size_t)n);
    const auto t0 = std::chrono::steady_clock::now();
    if (mlp_infer(m, pts, n, starts.data()) != 0) return -100;
    if (stats) stats[1] = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  int T = threads > 0 ? threads : cpu_share();
  T = std::max(1, std::min(T, n));
  std::vector<double> geo_ms(T, 0.0), solve_ms(T, 0.0);
  std::vector<int> bad(T, 0), fb(T, 0);
  auto work = [&](int id) {
    for (long long i = id; i < n; i += T) {
      const double *pt = pts + 4 * i;
      const double a = pt[0], p = pt[1], e = pt[2], iota = pt[3];
      double *x = maps + 3 * i;
      const auto t0 = std::chrono::steady_clock::now();
      Geodesic<HostCx> g;
      double Om[3];
      const bool ok = g.init(p, e, cos(iota), a, 1) == 0;
      if (ok) g.frequencies(Om);
      const auto t1 = std::chrono::steady_clock::now();
      geo_ms[id] += std::chrono::duration<double, std::milli>(t1 - t0).count();
      if (!ok || !std::isfinite(Om[0]) || !std::isfinite(Om[1]) || !std::isfinite(Om[2])) {
        x[0] = x[1] = x[2] = std::nan("");
        if (iters) iters[i] = 0;
        if (resid) resid[i] = std::nan("");
        bad[id]++;
        continue;
      }
      FreqMap fm{Om, 1.0, e, iota};
      int it;
      if (mode == 0) {
        x[0] = 1. / sqrt(p);
        x[1] = 1.0;
        x[2] = a;
        it = fm.solve(x, tol);
      } else {
        for (int k = 0; k < 3; k++) x[k] = starts[3 * i + k];
        it = fm.polish(x, tol, 8);
        if (it < 0 || !FreqMap::near_start(x, starts.data() + 3 * i, 0.05)) {
          x[0] = 1. / sqrt(p);
          x[1] = 1.0;
          x[2] = a;
          it = -fm.solve(x, tol);
          fb[id]++;
        }
      }
      solve_ms[id] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
      double f[3];
      fm.residual(x, f);
      const double r = fabs(f[0]) + fabs(f[1]) + fabs(f[2]);
      if (iters) iters[i] = it;
      if (resid) resid[i] = r;
      if (!std::isfinite(r)) bad[id]++;
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; t++) pool.emplace_back(work, t);
  work(0);
  for (auto &t : pool) t.join();
  int nbad = 0;
  for (int t = 0; t < T; t++) {
    nbad += bad[t];
    if (stats) {
      stats[0] += geo_ms[t];
      stats[2] += solve_ms[t];
      stats[3] += fb[t];
    }
  }
  return nbad;
}

int aak_set_tuning(aak_handle h, int feed_depth, int expansion_shift) {
  if (!h) return fail(-1, "null handle");
  if (feed_depth == 0 || feed_depth > 256) return fail(-2, "feed depth must be in [1, 256] (or < 0: keep)");
  if (expansion_shift > 3) return fail(-2, "expansion shift must be in [0, 3] (or < 0: keep)");
  if (feed_depth > 0) {
    std::lock_guard<std::mutex> lk(h->mu);
    h->feed_depth = feed_depth;
  }
  if (expansion_shift >= 0) h->cfg.exp_shift = expansion_shift;
  if (expansion_shift == -2) h->cfg.exp_shift = -1;  // back to the precision's default
  return 0;
}

int aak_set_nmax(aak_handle h, int n_max) {
  if (!h) return fail(-1, "null handle");
  if (n_max > 200) return fail(-2, "n_max above 200");
  h->cfg.n_max = n_max > 0 ? n_max : 0;
  return 0;
}

int aak_set_precision(aak_handle h, int mode) {
  if (!h) return fail(-1, "null handle");
  if (mode != 0 && mode != 1) return fail(-2, "precision mode must be 0 (FP64) or 1 (mixed)");
  h->cfg.mixed = mode;
  return 0;
}

int aak_set_lanes(aak_handle h, int n) {
  if (!h) return fail(-1, "null handle");
  if (n < 1 || n > kMaxLanes) return fail(-2, "lanes must be in [1, " + std::to_string(kMaxLanes) + "]");
  h->nlanes = n;
  return 0;
}

int aak_set_host_threads(aak_handle h, int n) {
  if (!h) return fail(-1, "null handle");
  if (n < 0 || n > 256) return fail(-2, "host threads must be in [0, 256]");
  stop_workers(h);
  h->host_threads = n;
  return 0;
}

int aak_input_data(aak_handle h, const double *hI_f, const double *hII_f, const double *ASDinv1,
                   const double *ASDinv2, int len) {
  if (!h) return fail(-1, "null handle");
  if (len != h->nbins) return fail(-2, "input_data: len must equal length/2+1");
  const size_t nb = (size_t)h->nbins;
  hipStream_t st = h->lane[0].stream;
  if (h->fused_fft) {
    // bin k = k1 + R k2 (k < M) goes to k1 C + k2, the Nyquist bin stays at M
    const FftArgs &F = h->fft;
    std::vector<double2> d(2 * nb);
    std::vector<double> w(2 * nb);
    const double2 *src_d[2] = {reinterpret_cast<const double2 *>(hI_f), reinterpret_cast<const double2 *>(hII_f)};
    const double *src_w[2] = {ASDinv1, ASDinv2};
    for (int ch = 0; ch < 2; ch++) {
      double2 *dd = d.data() + ch * nb;
      double *ww = w.data() + ch * nb;
      for (int k1 = 0; k1 < F.R; k1++)
        for (int k2 = 0; k2 < F.C; k2++) {
          dd[(size_t)k1 * F.C + k2] = src_d[ch][k1 + (size_t)F.R * k2];
          ww[(size_t)k1 * F.C + k2] = src_w[ch][k1 + (size_t)F.R * k2];
        }
      dd[F.M] = src_d[ch][F.M];
      ww[F.M] = src_w[ch][F.M];
    }
    HIPCHK(hipMemcpyAsync(h->d_data, d.data(), sizeof(double2) * 2 * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->d_w, w.data(), sizeof(double) * 2 * nb, hipMemcpyHostToDevice, st));
  } else {
    HIPCHK(hipMemcpyAsync(h->d_data, hI_f, sizeof(double2) * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->d_data + nb, hII_f, sizeof(double2) * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->d_w, ASDinv1, sizeof(double) * nb, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(h->d_w + nb, ASDinv2, sizeof(double) * nb, hipMemcpyHostToDevice, st));
  }
  HIPCHK(hipStreamSynchronize(st));
  h->have_data = 1;
  return 0;
}

int aak_gpu_gen_AAK(aak_handle h, double iota, double s, double p, double e, double M, double mu, double gamma,
                    double psi, double alph, double theta_S, double phi_S, double theta_K, double phi_K, double D) {
  if (!h) return fail(-1, "null handle");
  SourceParams sp{iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D};
  WaveformPlan &plan = h->single;
  std::string err;
  auto t0 = std::chrono::steady_clock::now();
  // the staging buffer is free: every call ends with the compute stream idle,
  // and the compute stream waited for the upload
  PinnedSink sink(&h->single_pin);
  h->have_traj = 0;
  int rc = plan_safe(h->cfg, sp, &plan, &err, &sink);
  h->host_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (rc < 0) {
    h->state = 0;
    return fail(rc, err);
  }
  h->have_traj = 1;
  Lane &L = h->lane[0];
  rc = enqueue_waveform(h, L, plan.P, plan.tab, plan.tab_n, 0, nullptr);
  if (rc) return rc;
  int flags[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  HIPCHK(hipMemcpyAsync(flags, L.d_flags, sizeof(int) * 8, hipMemcpyDeviceToHost, L.stream));
  HIPCHK(hipStreamSynchronize(L.stream));
  read_stage_times(h, 0, 3);
  if (flags[3]) {
    h->state = 0;
    return fail(-3, "non-positive radial frequency at t0");
  }
  h->state = 1;
  const DenseParams &P = plan.P;
  h->gen_Mphys = M;
  h->gen_coslam = P.coslam;
  h->gen_ip = P.ip;
  int *I = h->info;
  I[0] = plan.traj.steps;
  I[1] = P.nmodes;
  I[2] = P.ip;
  I[3] = P.ib;
  I[4] = P.K;
  I[5] = P.i0;
  I[6] = flags[1];
  I[7] = flags[0];
  I[8] = plan.knots.j0;
  I[9] = plan.knots.plunged;
  I[10] = flags[2];
  I[11] = flags[4];
  return 0;
}

int aak_run_phase_trajectory(aak_handle h, double iota, double s, double p, double e, double M, double mu,
                             double gamma, double psi, double alph, double theta_S, double phi_S, double theta_K,
                             double phi_K, double D) {
  if (!h) return fail(-1, "null handle");
  SourceParams sp{iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D};
  std::string err;
  PinnedSink sink(&h->single_pin);
  h->have_traj = 0;
  const int rc = plan_safe(h->cfg, sp, &h->single, &err, &sink);
  if (rc < 0) return fail(rc, err);
  h->have_traj = 1;
  return 0;
}

int aak_get_trajectory(aak_handle h, double *t_in, double *e_in, double *v_in, double *M_in, double *S_in, int cap,
                       int *i_plunge, int *i_buffer, double *t_clip, double *interp_timestep, int *init_length) {
  if (!h) return fail(-1, "null handle");
  if (!h->have_traj) return fail(-2, "no trajectory: call run_phase_trajectory or gpu_gen_AAK first");
  if (!t_in || !e_in || !v_in || !M_in || !S_in || !i_plunge || !i_buffer || !t_clip || !interp_timestep ||
      !init_length)
    return fail(-2, "null argument");
  const WaveformPlan &plan = h->single;
  try {
    return knot_outputs(h->cfg, plan.knots, plan.traj, plan.par, t_in, e_in, v_in, M_in, S_in, cap, i_plunge,
                        i_buffer, t_clip, interp_timestep, init_length);
  } catch (const std::bad_alloc &) {
    return fail(-9, "host memory exhausted by the knot table");
  }
}

int aak_gen_phases(aak_handle h, const double *params, double *t, double *phase_r, double *phase_theta,
                   double *phase_phi, double *omega_r, double *omega_theta, double *omega_phi, double *ecc) {
  if (!h) return fail(-1, "null handle");
  if (!params) return fail(-2, "null parameters");
  SourceParams sp;
  std::memcpy(&sp, params, sizeof(SourceParams));
  WaveformPlan &plan = h->single;
  std::string err;
  PinnedSink sink(&h->single_pin);
  h->have_traj = 0;
  int rc = plan_safe(h->cfg, sp, &plan, &err, &sink);
  if (rc < 0) {
    h->state = 0;
    return fail(rc, err);
  }
  h->have_traj = 1;
  if (!h->d_phase) HIPCHK(hipMalloc(&h->d_phase, sizeof(double) * 8 * (size_t)h->N));
  // the phase output applies the Phidot guard to every forward sample (AAKPhase.cc:233-241)
  DenseParams P = plan.P;
  P.guard_last = P.N - 1;
  Lane &L = h->lane[0];
  rc = enqueue_waveform(h, L, plan.P, plan.tab, plan.tab_n, 0, nullptr, &P);
  if (rc) return rc;
  const DenseBuffers B = h->bufs(L, 0);
  launch_phase_out(P, B, h->d_phase, L.stream);
  HIPCHK(hipGetLastError());
  double *dst[8] = {t, phase_r, phase_theta, phase_phi, omega_r, omega_theta, omega_phi, ecc};
  for (int a = 0; a < 8; a++)
    if (dst[a])
      HIPCHK(hipMemcpyAsync(dst[a], h->d_phase + (size_t)a * h->N, sizeof(double) * h->N, hipMemcpyDeviceToHost,
                            L.stream));
  HIPCHK(hipStreamSynchronize(L.stream));
  h->state = 0;
  return 0;
}

int aak_likelihood(aak_handle h, double *like_out) {
  if (!h) return fail(-1, "null handle");
  if (!h->have_data) return fail(-2, "Likelihood before input_data");
  if (h->state != 1) return fail(-2, "Likelihood needs a waveform from gpu_gen_AAK");
  Lane &L = h->lane[0];
  h->mark(L, 3);
  int rc = enqueue_likelihood(h, L, L.d_like, true, L.d_flags);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(like_out, L.d_like, sizeof(double) * 2, hipMemcpyDeviceToHost, L.stream));
  HIPCHK(hipStreamSynchronize(L.stream));
  read_stage_times(h, 3, 5);
  h->state = 2;
  return 0;
}

int aak_likelihood_batch(aak_handle h, const double *params, int n, double *like_out, int *status) {
  return aak_likelihood_batch_nmax(h, params, nullptr, n, like_out, status);
}

int aak_likelihood_batch_nmax(aak_handle h, const double *params, const int *nmax, int n, double *like_out,
                              int *status) {
  if (!h) return fail(-1, "null handle");
  if (!h->have_data) return fail(-2, "likelihood batch before input_data");
  if (n < 0 || (n > 0 && (!params || !like_out))) return fail(-2, "bad batch arguments");
  if (n == 0) return 0;
  if (nmax)
    for (int i = 0; i < n; i++)
      if (nmax[i] < 0 || nmax[i] > 200) return fail(-2, "per-template harmonic count outside 0..200");
  const auto t_call = std::chrono::steady_clock::now();
  int rc = start_pool(h);
  if (rc) return rc;
  const int nl = h->nlanes;
  for (int l = 1; l < nl; l++) {
    rc = alloc_lane(h, h->lane[l]);
    if (rc) return rc;
  }

  if ((size_t)n * 2 > h->like_batch_cap) {
    if (h->d_like_batch) HIPCHK(hipFree(h->d_like_batch));
    h->d_like_batch = nullptr;
    h->like_batch_cap = 0;
    HIPCHK(hipMalloc(&h->d_like_batch, sizeof(double) * 2 * (size_t)n));
    h->like_batch_cap = (size_t)n * 2;
  }
  if (h->like_host.reserve((size_t)n * 2)) return fail(-100, "pinned host allocation failed");
  std::vector<int> st((size_t)n, 0);
  const int prof = h->profiling;
  if (prof && h->batch_ev.size() < (size_t)6 * n) {
    const size_t old = h->batch_ev.size();
    h->batch_ev.resize((size_t)6 * n, nullptr);
    for (size_t k = old; k < h->batch_ev.size(); k++) HIPCHK(hipEventCreate(&h->batch_ev[k]));
  }
  // Order by estimated host cost: the host stage costs a few ms plus O(K)
  // with K, the knot count, ~ 99 max(1, t_rr / T_fit) where
  // t_rr = length dt mu / (M^2 M_sun) is the observation time in
  // radiation-reaction units (AAKpy.cc:24-26).  Workers take templates from
  // the cheap end while fewer than kFeedDepth plans are ready or in progress
  // from there (a steady supply for the GPU from the first millisecond), and
  // from the expensive end otherwise (the longest plans, which bound the
  // batch, start early).
  std::vector<std::pair<double, long long>> cost((size_t)n);
  for (long long i = 0; i < n; i++) {
    const double M = params[14 * i + 4], mu = params[14 * i + 5];
    const double tr = h->cfg.length * h->cfg.dt / kMsunSec / M / M * mu;
    const double c = (tr > h->cfg.T_fit) ? tr / h->cfg.T_fit : 1.0;
    cost[(size_t)i] = {std::isfinite(c) ? c : 0.0, i};
  }
  std::stable_sort(cost.begin(), cost.end(),
                   [](const std::pair<double, long long> &a, const std::pair<double, long long> &b) {
                     return a.first > b.first;
                   });
  {
    std::lock_guard<std::mutex> lk(h->mu);
    h->b_order.resize((size_t)n);
    for (long long i = 0; i < n; i++) h->b_order[(size_t)i] = cost[(size_t)i].second;
    h->free_slots.clear();
    h->ready_slots.clear();
    for (int s = (int)h->ring.size() - 1; s >= 0; s--) h->free_slots.push_back(s);
    h->b_params = params;
    h->b_nmax = nmax;
    h->b_n = n;
    h->b_next = 0;
    h->b_end = n;
    h->cheap_inflight = 0;
  }
  h->cv_work.notify_all();
  int err_rc = 0;
  std::string err_msg;
  double host_sum = 0.0, host_max = 0.0, issue_ms = 0.0, wait_ms = 0.0;
  for (long long done = 0; done < n; done++) {
    int sidx;
    const auto t_wait = std::chrono::steady_clock::now();
    {
      std::unique_lock<std::mutex> lk(h->mu);
      h->cv_ready.wait(lk, [&] { return !h->ready_slots.empty(); });
      sidx = h->ready_slots.front();
      h->ready_slots.pop_front();
    }
    const auto t_issue = std::chrono::steady_clock::now();
    wait_ms += std::chrono::duration<double, std::milli>(t_issue - t_wait).count();
    RingSlot &slot = *h->ring[(size_t)sidx];
    const long long i = slot.tpl;
    st[i] = slot.rc < 0 ? slot.rc : 0;
    host_sum += slot.plan_ms;
    host_max = std::max(host_max, slot.plan_ms);
    if (st[i] == 0 && err_rc == 0) {
      if (prof) h->evs = h->batch_ev.data() + 6 * i;
      // templates alternate over the lanes; each lane alternates its two table slots
      Lane &L = h->lane[done % nl];
      slot.P.full_dense = 0;  // likelihoods read e, the rates, phases and amplitude up to the plunge
      rc = enqueue_waveform(h, L, slot.P, slot.pin.p, slot.tab_n, (int)((done / nl) & 1), slot.copied);
      if (!rc) {
        slot.copy_pending = true;
        rc = enqueue_likelihood(h, L, h->d_like_batch + 2 * i, false, L.d_flags);
      }
      if (rc) {
        err_rc = rc;
        err_msg = g_err;
      }
    }
    if (err_rc && st[i] == 0) st[i] = err_rc;
    {
      std::lock_guard<std::mutex> lk(h->mu);
      h->free_slots.push_back(sidx);
    }
    h->cv_work.notify_one();
    issue_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_issue).count();
  }
  {
    std::lock_guard<std::mutex> lk(h->mu);
    h->b_params = nullptr;
    h->b_nmax = nullptr;
  }
  h->evs = h->ev;
  h->state = 0;
  if (err_rc) {
    for (int l = 0; l < nl; l++) (void)hipStreamSynchronize(h->lane[l].stream);
    (void)hipStreamSynchronize(h->cstream);
    if (status) std::memcpy(status, st.data(), sizeof(int) * (size_t)n);
    return fail(err_rc, err_msg);
  }
  for (int l = 1; l < nl; l++) HIPCHK(hipStreamSynchronize(h->lane[l].stream));
  Lane &L0 = h->lane[0];
  HIPCHK(hipMemcpyAsync(h->like_host.p, h->d_like_batch, sizeof(double) * 2 * (size_t)n, hipMemcpyDeviceToHost,
                        L0.stream));
  HIPCHK(hipStreamSynchronize(L0.stream));
  HIPCHK(hipStreamSynchronize(h->cstream));
  {
    // with the pool idle and every upload done, bring each slot to the largest
    // table planned so far, so later batches do not allocate pinned memory
    const size_t hint = h->pin_hint.load();
    for (auto &slot : h->ring)
      if (slot->pin.cap < hint && slot->pin.reserve(hint)) return fail(-100, "pinned host allocation failed");
    // and every worker's evolution arrays at the largest planned so far
    std::unique_lock<std::mutex> lk(h->mu);
    const size_t target = std::min(h->knot_hint, kIdleGrowCap);
    h->cv_ready.wait(lk, [&] {
      for (size_t c : h->worker_cap)
        if (c < target) return false;
      return true;
    });
  }
  h->batch_host_sum = host_sum;
  h->batch_host_max = host_max;
  h->batch_issue_ms = issue_ms;
  h->batch_wait_ms = wait_ms;
  h->batch_wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_call).count();
  if (prof) {
    for (int k = 0; k < 5; k++) h->batch_ms[k] = 0.0;
    float ms = 0.f;
    for (long long i = 0; i < n; i++) {
      if (st[i]) continue;
      const hipEvent_t *e = h->batch_ev.data() + 6 * i;
      for (int k = 0; k < 5; k++)
        if (hipEventElapsedTime(&ms, e[k], e[k + 1]) == hipSuccess) h->batch_ms[k] += ms;
    }
    h->batch_profiled = 1;
  }
  const double qnan = std::nan("");
  for (long long i = 0; i < n; i++) {
    // a NaN result of a planned template: rejected on the device (non-positive
    // radial rate at t = 0, as the single call reports it)
    if (st[i] == 0 && std::isnan(h->like_host.p[2 * i + 1])) st[i] = -3;
    like_out[2 * i] = st[i] ? qnan : h->like_host.p[2 * i];
    like_out[2 * i + 1] = st[i] ? qnan : h->like_host.p[2 * i + 1];
  }
  if (status) std::memcpy(status, st.data(), sizeof(int) * (size_t)n);
  return 0;
}

int aak_get_waveform(aak_handle h, double *t, double *hI, double *hII) {
  if (!h) return fail(-1, "null handle");
  if (h->state == 0) return fail(-2, "GetWaveform before gpu_gen_AAK");
  const size_t n = (size_t)h->N + 2;
  const Lane &L = h->lane[0];
  HIPCHK(hipMemcpyAsync(t, L.d_t, sizeof(double) * n, hipMemcpyDeviceToHost, L.stream));
  if (h->state == 2) {
    HIPCHK(hipMemcpyAsync(hI, L.d_hf, sizeof(double2) * h->nbins, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(hipMemcpyAsync(hII, L.d_hf + h->nbins, sizeof(double2) * h->nbins, hipMemcpyDeviceToHost, L.stream));
  } else {
    HIPCHK(hipMemcpyAsync(hI, L.d_h, sizeof(double) * n, hipMemcpyDeviceToHost, L.stream));
    HIPCHK(hipMemcpyAsync(hII, L.d_h + h_stride(h->N), sizeof(double) * n, hipMemcpyDeviceToHost, L.stream));
  }
  HIPCHK(hipStreamSynchronize(L.stream));
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

int aak_get_batch_timings(aak_handle h, double *ms, int n) {
  if (!h) return fail(-1, "null handle");
  if (!h->batch_profiled) return fail(-2, "no profiled batch");
  const int m = n < 5 ? n : 5;
  for (int i = 0; i < m; i++) ms[i] = h->batch_ms[i];
  return m;
}

int aak_get_batch_host_ms(aak_handle h, double *sum_ms, double *max_ms) {
  if (!h) return fail(-1, "null handle");
  if (sum_ms) *sum_ms = h->batch_host_sum;
  if (max_ms) *max_ms = h->batch_host_max;
  return 0;
}

int aak_get_pool_stats(aak_handle h, double *out, int n) {
  if (!h) return fail(-1, "null handle");
  const double threads = (double)h->workers.size();
  const double busy = (threads > 0 && h->batch_wall_ms > 0) ? h->batch_host_sum / (threads * h->batch_wall_ms) : 0.0;
  const double v[7] = {threads, h->batch_host_sum, h->batch_host_max, h->batch_wall_ms, busy, h->batch_issue_ms,
                       h->batch_wait_ms};
  const int m = n < 7 ? n : 7;
  for (int i = 0; i < m; i++) out[i] = v[i];
  return m;
}

int aak_get_info(aak_handle h, int *info, int n) {
  if (!h) return fail(-1, "null handle");
  const int m = n < 12 ? n : 12;
  for (int i = 0; i < m; i++) info[i] = h->info[i];
  return m;
}

// ---------------------------------------------------------------- test hooks
static int test_plan(const double *params, double T_fit, int length, double dt, int LISA, int backint, int n_max,
                     WaveformPlan *plan) {
  WaveConfig cfg;
  cfg.T_fit = T_fit;
  cfg.length = length;
  cfg.dt = dt;
  cfg.LISA = LISA;
  cfg.backint = backint;
  cfg.n_max = n_max;
  SourceParams sp;
  std::memcpy(&sp, params, sizeof(double) * 14);
  std::string err;
  const int rc = plan_safe(cfg, sp, plan, &err);
  return rc < 0 ? fail(rc, err) : 0;
}

int aak_test_host_plan(const double *params, double T_fit, int length, double dt, int LISA, int backint, int n_max,
                       double *traj_out, double *knots_out, int cap, int *info, double *dparams) {
  WaveformPlan plan;
  const int rc = test_plan(params, T_fit, length, dt, LISA, backint, n_max, &plan);
  if (rc) return rc;
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
  const double *src[5] = {ks.kt(), ks.ky(0), ks.ky(1), ks.ky(2), ks.ky(3)};
  for (int a = 0; a < 5; a++) std::memcpy(knots_out + (size_t)a * cap, src[a], sizeof(double) * ks.K);
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

int aak_test_knot_table(const double *params, double T_fit, int length, double dt, int LISA, int backint,
                        double *table, int cap, int *grid) {
  WaveformPlan plan;
  const int rc = test_plan(params, T_fit, length, dt, LISA, backint, 0, &plan);
  if (rc) return rc;
  if (plan.tab_n > (size_t)cap) return fail(-6, "table buffer too small");
  std::memcpy(table, plan.tab, sizeof(double) * plan.tab_n);
  grid[0] = plan.P.grid.K;
  grid[1] = plan.P.grid.head_off;
  grid[2] = plan.P.grid.steps;
  grid[3] = plan.P.grid.flat;
  return (int)plan.tab_n;
}

int aak_test_knot_eval(const double *params, double T_fit, int length, double dt, int LISA, int backint,
                       const double *t, int n, int *k_out, double *y_out) {
  WaveformPlan plan;
  const int rc = test_plan(params, T_fit, length, dt, LISA, backint, 0, &plan);
  if (rc) return rc;
  const double *head = plan.tab + (size_t)kTableArrays * plan.P.grid.K;
  for (int i = 0; i < n; i++) {
    k_out[i] = knot_index(plan.P.grid, head, t[i]);
    for (int a = 0; a < 4; a++) y_out[4 * (size_t)i + a] = table_eval(plan.tab, plan.P.grid, a, t[i]);
  }
  return 0;
}

int aak_test_phidot_guard(double *rate, int n, int i0, int guard_last) {
  if (!rate || n < 1 || i0 < 0 || guard_last >= n || i0 > guard_last) return fail(-2, "guard range outside the array");
  // the count and the first non-positive sample, as k_rates records them
  int count = 0, first = -1;
  for (int i = i0 + 1; i <= guard_last; i++)
    if (rate[i] <= 0.) {
      if (first < 0) first = i;
      count++;
    }
  int flags[8] = {count, 0, 0, 0, 0, first < 0 ? 0 : 0x7fffffff - first, 0, 0};
  int gw[kGuardGrid + 1];
  for (int k = 0; k < kGuardGrid; k++) gw[k] = -2;
  gw[kGuardGrid] = 0;
  double *d_rate = nullptr;
  int *d_flags = nullptr, *d_gw = nullptr;
  auto release = [&]() {
    if (d_rate) (void)hipFree(d_rate);
    if (d_flags) (void)hipFree(d_flags);
    if (d_gw) (void)hipFree(d_gw);
  };
  if (hipMalloc(&d_rate, sizeof(double) * n) != hipSuccess || hipMalloc(&d_flags, sizeof(flags)) != hipSuccess ||
      hipMalloc(&d_gw, sizeof(gw)) != hipSuccess) {
    release();
    return fail(-6, "device allocation failed");
  }
  DenseParams P;
  std::memset(&P, 0, sizeof(P));
  P.N = n - 1;
  P.i0 = i0;
  P.guard_last = guard_last;
  DenseBuffers B;
  std::memset(&B, 0, sizeof(B));
  B.rate[1] = d_rate;
  B.flags = d_flags;
  B.guard_work = d_gw;
  hipError_t st = hipMemcpy(d_rate, rate, sizeof(double) * n, hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(d_flags, flags, sizeof(flags), hipMemcpyHostToDevice);
  if (st == hipSuccess) st = hipMemcpy(d_gw, gw, sizeof(gw), hipMemcpyHostToDevice);
  if (st == hipSuccess) {
    launch_phidot_guard(P, B, FoldReset{}, 0);
    st = hipGetLastError();
  }
  if (st == hipSuccess) st = hipMemcpy(rate, d_rate, sizeof(double) * n, hipMemcpyDeviceToHost);
  release();
  if (st != hipSuccess) return fail(-6, hipGetErrorString(st));
  return count;
}

int aak_test_bessel_five(int n, double e, double e_bound, double *J) {
  if (n < 1 || n > kMaxModes) return fail(-2, "harmonic outside 1..128");
  if (!(e >= 0.0) || e > e_bound || e_bound > kSeriesMaxE) return fail(-2, "eccentricity outside the series range");
  static std::vector<double> tab;
  static std::once_flag once;
  std::call_once(once, []() {
    const size_t nc = (size_t)(kBesselKMax + 1) * kBesselMMax;
    tab.resize(nc + kBesselKMax + 1);
    bessel_series_table(tab.data());
    bessel_log_factorials(tab.data() + nc);
  });
  unsigned char terms[kMaxModes + 1];
  bessel_series_terms(e_bound, n, terms);
  bessel_five_series(n, e, tab.data(), tab.data() + (size_t)(kBesselKMax + 1) * kBesselMMax, terms[n], terms[0], J);
  return terms[n];
}

int aak_test_taylor_amplitudes(int n_max, int n, double e_ref, double del, double d, double *out) {
  if (n < 1 || n > n_max || n_max > kMaxModes) return fail(-2, "harmonic outside 1..n_max (<= 128)");
  if (!(e_ref >= 0.0) || !(del >= 0.0) || e_ref + del > kSeriesMaxE) return fail(-2, "eccentricity range");
  static std::vector<double> tab;
  static std::once_flag once;
  std::call_once(once, []() {
    const size_t nc = (size_t)(kBesselKMax + 1) * kBesselMMax;
    tab.resize(nc + kBesselKMax + 1);
    bessel_series_table(tab.data());
    bessel_log_factorials(tab.data() + nc);
  });
  const double *ctab = tab.data(), *lgf = tab.data() + (size_t)(kBesselKMax + 1) * kBesselMMax;
  unsigned char terms[kMaxModes + 1];
  bessel_series_terms(e_ref + del, n_max, terms);
  const int p = taylor_order(n_max, del);
  if (p < 0) return fail(-2, "no Taylor order for this range");
  double sq[kTaylorMax + 1], coef[3 * (kTaylorMax + 1)];
  taylor_sqrt1me2(e_ref, p, sq);
  taylor_amplitudes(n, p, e_ref, sq, ctab, lgf, terms[n], coef);
  taylor_eval(coef, p, d, out);
  // direct evaluation at e = e_ref + d
  const double e = e_ref + d;
  double J[5];
  bessel_five_series(n, e, ctab, lgf, terms[n], terms[0], J);
  out[3] = -n * (J[0] - 2. * e * J[1] + 2. / n * J[2] + 2. * e * J[3] - J[4]);
  out[4] = -n * sqrt(1 - e * e) * (J[0] - 2. * J[2] + J[4]);
  out[5] = 2. * J[2];
  return p;
}

int aak_test_inspiral_rates(double p, double e, double cosi, double a, double *out) {
  return inspiral_rates<HostCx>(p, e, cosi, a, out);
}

int aak_test_inner_product(aak_handle h, const double *hI, const double *hII, double *like, double *spec,
                           int *fused) {
  if (!h) return fail(-1, "null handle");
  if (!h->have_data) return fail(-2, "inner product before input_data");
  const size_t n = (size_t)h->N, ld = h_stride(h->N);
  Lane &L = h->lane[0];
  HIPCHK(hipMemcpyAsync(L.d_h, hI, sizeof(double) * n, hipMemcpyHostToDevice, L.stream));
  HIPCHK(hipMemcpyAsync(L.d_h + ld, hII, sizeof(double) * n, hipMemcpyHostToDevice, L.stream));
  const int rc = enqueue_likelihood(h, L, L.d_like, spec != nullptr, nullptr);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(like, L.d_like, sizeof(double) * 2, hipMemcpyDeviceToHost, L.stream));
  if (spec) HIPCHK(hipMemcpyAsync(spec, L.d_hf, sizeof(double2) * 2 * h->nbins, hipMemcpyDeviceToHost, L.stream));
  HIPCHK(hipStreamSynchronize(L.stream));
  h->state = 0;  // the time-domain buffers no longer hold a generated waveform
  if (fused) *fused = h->fused_fft ? 1 : 0;
  return 0;
}

int aak_test_force_fold_replay(aak_handle h, int on) {
  if (!h) return fail(-1, "null handle");
  h->force_fold_replay = on == 2 ? 2 : (on ? 1 : 0);
  return 0;
}

int aak_test_fold_mode(aak_handle h, int mode) {
  if (!h) return fail(-1, "null handle");
  if (mode != 0 && mode != 1) return fail(-2, "fold mode must be 0 or 1");
  h->fold_mode = mode;
  return 0;
}

int aak_get_inspiral(aak_handle h, double *t, double *p, double *e, int cap) {
  if (!h) return fail(-1, "null handle");
  if (h->state == 0) return fail(-2, "no waveform generated");
  if (!t || !p || !e || cap < 1) return fail(-2, "bad arguments");
  const int n = std::min(h->gen_ip + 1, cap);
  const size_t L = (size_t)h->N + 1;
  std::vector<double> d(4 * L), tv(h->N);
  const Lane &ln = h->lane[0];
  // e, v, M, S are the first four dense arrays
  HIPCHK(hipMemcpyAsync(d.data(), ln.d_dense, sizeof(double) * 4 * L, hipMemcpyDeviceToHost, ln.stream));
  HIPCHK(hipMemcpyAsync(tv.data(), ln.d_t, sizeof(double) * h->N, hipMemcpyDeviceToHost, ln.stream));
  HIPCHK(hipStreamSynchronize(ln.stream));
  for (int i = 0; i < n; i++) {
    const double ee = d[i], v = d[L + i], M = d[2 * L + i], S = d[3 * L + i];
    t[i] = tv[i];
    p[i] = (1. - ee * ee) * pow(omega_phi(v, ee, h->gen_coslam, S, M) * h->gen_Mphys * kMsunSec, -2. / 3.);
    e[i] = ee;
  }
  return n;
}

int aak_kerr_constants(double p, double e, double iota, double a, double *ELQ) {
  if (!ELQ) return fail(-2, "null output");
  Geodesic<HostCx> g;
  if (g.init(p, e, cos(iota), a, 1)) return fail(-5, "geodesic constants failed");
  ELQ[0] = g.E;
  ELQ[1] = g.Lz;
  ELQ[2] = g.Q;
  return 0;
}

int aak_test_get_dense(aak_handle h, double *out) {
  if (!h) return fail(-1, "null handle");
  if (h->state == 0) return fail(-2, "no waveform generated");
  const size_t L = (size_t)h->N + 1;
  const Lane &ln = h->lane[0];
  HIPCHK(hipMemcpyAsync(out, ln.d_dense, sizeof(double) * 10 * L, hipMemcpyDeviceToHost, ln.stream));
  HIPCHK(hipStreamSynchronize(ln.stream));
  return 0;
}

}  // extern "C"
