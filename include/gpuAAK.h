/*
 * gpuAAK.h -- C ABI of the MI355X AAK waveform + LISA likelihood library
 * (emri_kludge_suite_amd/libaak_hip.so).
 *
 * The entry points are what a binding of the reference's GPU path needs: the
 * GPUAAK class of include/AAK_manager.hh (constructor, input_data,
 * gpu_gen_AAK, Likelihood, GetWaveform, destructor; bound by the Cython
 * pygpuAAK/GPUAAK.pyx:50-93) and the two free functions of include/gpuAAK.h
 * (OmegaPhi, PNevolution).  All functions are plain C: pointers, sizes and
 * status codes; no exceptions cross the boundary.  Host arrays are copied,
 * never retained.  Calls on one handle must not overlap; separate handles are
 * independent.  Status: 0 = success, < 0 = error (message via aak_last_error).
 *
 * Semantics follow the reference CPU AAK (AAK.cc): natural spline, Kepler
 * mean anomaly, unflipped theta_S, Phidot guard and the reference's LISA=false
 * response.
 */
#ifndef EMRI_AAK_GPUAAK_H
#define EMRI_AAK_GPUAAK_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aak_handle_s *aak_handle;

/* GPUAAK::GPUAAK(T_fit, init_length, length, init_dt, dt, LISA, backint)
 * (AAK_manager.hh:106-112, AAK_manager.cu:42-133).  init_length/init_dt are
 * accepted for signature compatibility; knot storage is sized per waveform. */
int aak_create(double T_fit, int init_length, int length, double init_dt, double dt,
               int LISA, int backint, aak_handle *out);

/* GPUAAK::~GPUAAK (AAK_manager.cu:460-491). */
int aak_destroy(aak_handle h);

/* GPUAAK::input_data(hI_f, hII_f, ASDinv1, ASDinv2, len) (AAK_manager.cu:135-152).
 * hI_f/hII_f: len complex values as interleaved (re, im) doubles, already
 * multiplied by the weights; ASDinv*: len weights; len must be length/2+1. */
int aak_input_data(aak_handle h, const double *hI_f, const double *hII_f,
                   const double *ASDinv1, const double *ASDinv2, int len);

/* GPUAAK::gpu_gen_AAK(iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S,
 * theta_K, phi_K, D) (AAK_manager.cu:154-260): waveform into device memory. */
int aak_gpu_gen_AAK(aak_handle h, double iota, double s, double p, double e, double M,
                    double mu, double gamma, double psi, double alph, double theta_S,
                    double phi_S, double theta_K, double phi_K, double D);

/* GPUAAK::Likelihood(like_out[2]) (AAK_manager.cu:352-452):
 * like_out = {4 <h w, d w>, 4 <h w, h w>} over both channels. */
int aak_likelihood(aak_handle h, double *like_out);

/* GPUAAK::GetWaveform(t, hI, hII) (AAK_manager.cu:454-458): length+2 doubles
 * each; time domain after aak_gpu_gen_AAK, the weighted spectrum (interleaved)
 * after aak_likelihood, as in the reference. */
int aak_get_waveform(aak_handle h, double *t, double *hI, double *hII);

/* Batched likelihoods against the data of aak_input_data: params holds n
 * rows of the 14 gpu_gen_AAK parameters (reference order); like_out receives
 * n rows {4 <h w, d w>, 4 <h w, h w>}.  Replaces the per-template loop of
 * gpu_gen_AAK + Likelihood in the reference's caller (pygpuAAK.py:205-224,
 * GPUAAK.pyx:84-93 WaveformThroughLikelihood); each row equals that pair of
 * calls.  The host stages run on a pool of threads (aak_set_host_threads)
 * while the GPU works on earlier templates.  status (optional, n ints) gets
 * 0 or the negative code of a template whose host stage failed; its row is
 * NaN.  Returns 0, or < 0 if the batch as a whole failed.  The waveform
 * buffers are left in an unspecified state (call gpu_gen_AAK again before
 * GetWaveform). */
int aak_likelihood_batch(aak_handle h, const double *params, int n, double *like_out, int *status);

/* aak_likelihood_batch with a harmonic count per template (nmax: n ints,
 * 0 = the reference's rule, > 0 = n_max as aak_set_nmax sets it; NULL = the
 * handle's setting for all), for batches mixing eccentricities; no
 * counterpart in the reference. */
int aak_likelihood_batch_nmax(aak_handle h, const double *params, const int *nmax, int n, double *like_out,
                              int *status);

/* Arithmetic of the mode sum; no counterpart in the reference.  mode 0
 * (default): plain FP64, within relative L2 1e-10 of the reference CPU AAK.
 * mode 1: mixed precision -- phases, their reduction mod 2 pi, the Bessel
 * expansions and the likelihood sums stay FP64, the per-harmonic rotation,
 * amplitude and response arithmetic runs in FP32 (noise-weighted mismatch
 * against mode 0 below 1e-6).  Blocks outside the Taylor form (fast
 * evolution, e > 0.66) keep the FP64 per-harmonic evaluation. */
int aak_set_precision(aak_handle h, int mode);

/* Host threads of the batch pool (0 = AAK_HOST_THREADS, else the CPUs the
 * process may run on: its affinity mask, bounded by a cgroup CPU quota); no
 * counterpart in the reference. */
int aak_set_host_threads(aak_handle h, int n);

/* Host pool of the last batch; no counterpart in the reference.
 * out[0..6] = pool threads, host-stage ms summed over the templates, longest
 * template's host ms, wall ms of the batch call, busy fraction
 * (sum / (threads x wall)), ms the issuing thread spent enqueueing device
 * work, ms it spent waiting for finished plans.  Returns the number written. */
int aak_get_pool_stats(aak_handle h, double *out, int n);

/* Tuning of the batch pipeline (no counterpart in the reference; results do
 * not depend on it beyond the mixed form's FP32 rounding): feed_depth = short
 * host plans kept ready for the GPU (1..256; < 0 keeps it), expansion_shift =
 * mode-sum expansion blocks of 256 << shift samples (0..3; -1 keeps it, -2
 * restores the precision's default, 2 mixed / 0 FP64). */
int aak_set_tuning(aak_handle h, int feed_depth, int expansion_shift);

/* Concurrent streams ("lanes", 1..8) over which aak_likelihood_batch spreads
 * its templates, each with its own device buffers (default AAK_LANES or 2);
 * no counterpart in the reference.  Results do not depend on it. */
int aak_set_lanes(aak_handle h, int n);

/* AAK phases and frequencies of one source, on the handle's grid: the
 * reference's CPU AAKphase (AAKpy.cc:53-88, AAKPhase.cc:13-374; Python
 * AAKwrapper.phase, AAKwrapper.pyx:90-133).  params: the 14 gpu_gen_AAK
 * parameters.  Each output (length doubles, any may be NULL): t, phase_r =
 * Phi, phase_theta = Phi+gamma, phase_phi = Phi+gamma+alpha, omega_r = Phidot,
 * omega_theta = Phidot+gammadot, omega_phi = Phidot+gammadot+alphadot,
 * eccentricity.  The waveform buffers are left unspecified. */
int aak_gen_phases(aak_handle h, const double *params, double *t, double *phase_r, double *phase_theta,
                   double *phase_phi, double *omega_r, double *omega_theta, double *omega_phi, double *ecc);

/* OmegaPhi(v, e, cosiota, s, M) (gpuAAK.h:13, gpuAAK.cc:17-26). */
double aak_OmegaPhi(double v, double e, double cosiota, double s, double M);

/* PNevolution(t_in, e_in, v_in, M_in, S_in, timestep, vlength, par, e_traj,
 * v_map, M_phys, M_map, S_phys, S_map, dt_map, steps, i_plunge, i_buffer,
 * backint, t_clip, interp_timestep, init_length) (gpuAAK.h:15-18,
 * gpuAAK.cc:29-267): the AK evolution, fit corrections and plunge search of
 * the host stage.  Inputs as in the reference: par[12] (mu and M_map[0] in
 * seconds, s_map[0], e_traj[0], iota, gamma, Phi0, theta_S, phi_S, theta_K,
 * phi_K, alpha), the sparse trajectory of `steps` (3..100) points, M_phys (in
 * solar masses) and S_phys.  Outputs: the knots t_in, e_in, v_in, M_in, S_in
 * (caller-owned, `cap` entries each; -6 when the knots do not fit),
 * i_plunge (relative to the first sample of the waveform), i_buffer, t_clip,
 * interp_timestep (the uniform knot spacing) and init_length (the knot count).
 * The knots span the reference CPU's whole range (AAK.cc:220-226), not the
 * GPU reference's cut at j_max. */
int aak_PNevolution(double *t_in, double *e_in, double *v_in, double *M_in, double *S_in, int cap,
                    double timestep, int vlength, const double *par, const double *e_traj,
                    const double *v_map, double M_phys, const double *M_map, double S_phys,
                    const double *S_map, const double *dt_map, int steps, int *i_plunge, int *i_buffer,
                    int backint, double *t_clip, double *interp_timestep, int *init_length);

/* GPUAAK::run_phase_trajectory(14 parameters) (AAK_manager.hh:118-132,
 * AAK_manager.cu:263-342): the host stage alone (sparse inspiral, frequency
 * map, PNevolution), kept in the handle; aak_gpu_gen_AAK runs it itself.
 * aak_get_trajectory copies the kept knots and PNevolution outputs (same
 * meaning as in aak_PNevolution); no counterpart accessor exists in the
 * reference, whose knots stay in private members. */
int aak_run_phase_trajectory(aak_handle h, double iota, double s, double p, double e, double M,
                             double mu, double gamma, double psi, double alph, double theta_S,
                             double phi_S, double theta_K, double phi_K, double D);
int aak_get_trajectory(aak_handle h, double *t_in, double *e_in, double *v_in, double *M_in, double *S_in,
                       int cap, int *i_plunge, int *i_buffer, double *t_clip, double *interp_timestep,
                       int *init_length);

/* Harmonic-count override; no counterpart in the reference, whose count is
 * (int)(30 e0), 4 below e0 = 0.135 (AAK.cc:582-583).  n_max <= 0 restores it. */
int aak_set_nmax(aak_handle h, int n_max);

/* Diagnostics of the last aak_gpu_gen_AAK; no counterpart in the reference.
 * info[0..11] = steps, nmodes, i_plunge, i_buffer, knots K, i0, fold fallbacks,
 * non-positive Phidot count, j0, plunged, 256-sample blocks summed per sample,
 * expansion blocks expanded again per 256 samples.  Returns the number written. */
int aak_get_info(aak_handle h, int *info, int n);

/* Stage timing; no counterpart in the reference (its timers are commented out,
 * AAK_manager.cu:170-194).  aak_set_profiling(h, 1) records HIP events on the
 * handle's stream; aak_get_timings returns ms[0..5] = host stage (wall clock),
 * rates+spline, phase folds, mode sum, rFFT, weighted reduction. */
int aak_set_profiling(aak_handle h, int on);
int aak_get_timings(aak_handle h, double *ms, int n);

/* Stage time sums (ms) over the templates of the last batch run with
 * profiling on: rates+spline, guard+folds, mode sum, FFT pass (or rFFT),
 * second FFT pass with the weighted sums (or the reduction). */
int aak_get_batch_timings(aak_handle h, double *ms, int n);

/* Host-stage time of the last batch's templates (ms, on the host pool):
 * their sum and the longest one.  No counterpart in the reference. */
int aak_get_batch_host_ms(aak_handle h, double *sum_ms, double *max_ms);

/* ---------------------------------------------------------------------
 * Learned start point of the frequency map (no counterpart in the
 * reference, whose ParMap starts its hybrid solver at (1/sqrt(p), M, s):
 * KSParMap.cc:397-440).  A 4-128-128-128-3 GELU network with BF16 weights
 * and FP32 accumulation on the matrix cores predicts (v_map, M_map/M, s_map)
 * from (a, p, e, iota); a Newton polish takes it to the solver's tolerance
 * and the hybrid solver remains the fallback.  Trained on the GPU by the
 * library's own forward, backward and Adam kernels.
 * ------------------------------------------------------------------- */
typedef struct aak_mlp_s *aak_mlp;

/* New model with He-initialised weights (seed) on the current device. */
int aak_mlp_create(int seed, aak_mlp *out);
int aak_mlp_destroy(aak_mlp m);
/* params: the 34051 FP32 parameters W1[4][128] b1[128] W2[128][128] b2[128]
 * W3[128][128] b3[128] W4[128][3] b4[3] (W[k][n]: input k to output n);
 * norm (optional): 14 floats, feature centres[4] and scales[4] of
 * (a, 1/sqrt(p), e, cos iota), target means[3] and scales[3] of
 * (v_map - 1/sqrt(p), M_map/M - 1, s_map - a).  Setting resets Adam's step. */
int aak_mlp_set_params(aak_mlp m, const float *params, const float *norm);
int aak_mlp_set_norm(aak_mlp m, const float *norm);
/* Copies them out; returns the parameter count. */
int aak_mlp_get_params(aak_mlp m, float *params, float *norm);
/* Network outputs Y[n][3] for normalised inputs X[n][4] (no normalisation). */
int aak_mlp_forward(aak_mlp m, const float *X, int n, float *Y);
/* Gradient of the mean squared error over all n rows (one batch) with
 * respect to the 34051 parameters, and the loss; no update. */
int aak_mlp_gradient(aak_mlp m, const float *X, const float *T, int n, float *grad, double *loss);
/* Adam training on normalised inputs X[n][4] and targets T[n][3]: `epochs`
 * passes of n/batch shuffled batches (seed), learning rate decaying from lr0
 * to lr1 (cosine); loss_per_epoch (optional, epochs doubles) gets the last
 * batch's loss of each epoch. */
int aak_mlp_train(aak_mlp m, const float *X, const float *T, int n, int epochs, int batch, float lr0, float lr1,
                  int seed, double *loss_per_epoch);
/* Start points maps[n][3] = (v_map, M_map/M, s_map) at pts[n][4] =
 * (a, p, e, iota), through the model's normalisation. */
int aak_mlp_predict_maps(aak_mlp m, const double *pts, int n, double *maps);

/* Frequency maps of n points pts[n][4] = (a, p, e, iota) at M = 1 (the map
 * is linear in M), on `threads` host threads (0: the CPUs the process may run
 * on).  mode 0: the reference's hybrid solver from (1/sqrt(p), 1, a) down to
 * sum|f| <= tol; mode 1: start points from the model m (one batched
 * inference), Newton polish to tol, the hybrid solver from the reference's
 * start where the polish fails.  maps[n][3] = (v_map, M_map, s_map);
 * iters[n] (optional): solver iterations (mode 0) or Newton steps (mode 1,
 * negative: fell back); resid[n] (optional): the final sum|f|; stats
 * (optional, 4 doubles): host ms of the geodesic frequencies, of the model
 * (mode 1), of the solves, and the fallback count.  Returns the number of
 * points whose geodesic or map failed (their maps are NaN). */
int aak_freqmap_batch(const double *pts, int n, int mode, aak_mlp m, double tol, int threads, double *maps,
                      int *iters, double *resid, double *stats);

/* The reference's interpolated last-stable-orbit p at (a, e, iota)
 * (GKTrajFast.cc:20-130, P_lso), as the sparse inspiral uses it. */
double aak_p_lso(double a, double e, double iota);
/* The same at n points aei[n][3] = (a, e, iota) on `threads` host threads. */
int aak_p_lso_batch(const double *aei, int n, double *out, int threads);

/* Start point of the frequency map in the host stage of this handle: 0 the
 * reference's (default; results as the reference CPU AAK), 1 the model m
 * plus Newton polish to min(1e-6, 1e-12 M) with the hybrid solver as
 * fallback (maps then differ from the reference's within its tolerance).
 * The model must outlive its use by the handle. */
int aak_set_map_init(aak_handle h, int mode, aak_mlp m);

/* Inspiral of the last aak_gpu_gen_AAK up to its plunge sample, as the
 * reference's AAK_Waveform writes it to <path>_traj.dat (AAK.cc:355-366):
 * t, p = (1 - e^2) (M Omega_phi)^(-2/3) and e per sample, at most cap rows;
 * returns the row count.  No counterpart as an entry point in the reference. */
int aak_get_inspiral(aak_handle h, double *t, double *p, double *e, int cap);

/* Constants of motion (E, L_z, Q) of the Kerr geodesic (p, e, iota) about
 * spin a (IEKG.cc:19-159), as AAK_Waveform.cc:98-100 prints them. */
int aak_kerr_constants(double p, double e, double iota, double a, double *ELQ);

/* The likelihood's inner product on caller-owned device buffers, with no
 * device allocation.  No counterpart in the reference: for callers that
 * cannot allocate (code under HIP graph capture, serving loops).
 * hI, hII: the two channels' time series (length doubles each); hI_f,
 * hII_f: the weighted data, length/2+1 complex bins as interleaved
 * (re, im) doubles; ASDinv1/2: their weights; like_out: 2 doubles,
 * {4 <h w, d w>, 4 <h w, h w>} as aak_likelihood; stream: a hipStream_t or
 * NULL.  Limits: length a power of two in [2^8, 2^21] (else -2); scratch is
 * one static device array of about 80 MiB per GPU, held until the process
 * ends; the first call at a new length copies that length's FFT tables and
 * waits for the copy, later calls only enqueue kernels; calls that overlap
 * on one GPU share the scratch and race.  Series and spectra 16-byte
 * aligned. */
int aak_inner_product_noalloc(const double *hI, const double *hII, const double *hI_f, const double *hII_f,
                              const double *ASDinv1, const double *ASDinv2, int length, double *like_out,
                              void *stream);

/* Settings and parameters of the waveform generator AAK_Waveform (SetPar,
 * KSTools.h:21-51); booleans as 0/1, angles in radians, masses in solar
 * masses, D in Gpc. */
typedef struct {
  char path[1024];
  int backint, LISA, traj, SNR, timing;
  int length;
  double dt, p, T, f, T_fit;
  double mu, M, s, e, iota, gamma, psi;
  double theta_S, phi_S, theta_K, phi_K, alpha, D;
} aak_setpar;

/* LoadSetPar(par, filename) (KSTools.cc:26-114): 1 on success, 0 with the
 * reference's message in aak_last_error() otherwise.  dt < 0 becomes
 * T * 31536000 / length, p < 0 the p of the initial GW frequency f. */
int aak_load_setpar(const char *filename, aak_setpar *out);

/* eLISA_Noise(f) (KSTools.cc:116-125): analytic eLISA PSD in 1/Hz. */
double aak_elisa_noise(double f);

/* Message of the last failed call on this thread. */
const char *aak_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
