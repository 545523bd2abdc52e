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
 * out[0..4] = pool threads, host-stage ms summed over the templates, longest
 * template's host ms, wall ms of the batch call, busy fraction
 * (sum / (threads x wall)).  Returns the number written. */
int aak_get_pool_stats(aak_handle h, double *out, int n);

/* Concurrent streams ("lanes", 1..4) over which aak_likelihood_batch spreads
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
 * info[0..9] = steps, nmodes, i_plunge, i_buffer, knots K, i0, fold fallbacks,
 * non-positive Phidot count, j0, plunged.  Returns the number written. */
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

/* Message of the last failed call on this thread. */
const char *aak_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
