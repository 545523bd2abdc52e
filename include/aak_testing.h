/*
 * aak_testing.h -- host-only entry points of libaak_hip.so used by the test
 * suite to check the host half of the waveform (sparse inspiral, frequency
 * map, knots, spline table) without a GPU.  No counterpart in the reference;
 * callers are tests that compare against the CPU restatement.
 */
#ifndef EMRI_AAK_TESTING_H
#define EMRI_AAK_TESTING_H

#ifdef __cplusplus
extern "C" {
#endif

/* Runs the host stage for params[14] (reference order).  traj_out[5*100]:
 * e_traj | v_map | M_map | s_map | dt_map.  knots_out[5*cap]: t | e | v | M | S.
 * info[0..15]: steps, K, i0, j0, j_min, j_max, j_max_temp, j_start, j_end,
 * plunged, points, i_plunge, i_buffer, nmodes, last_fold, map iterations sum.
 * dparams[0..3]: Phi0, zeta, t_plunge, t_zero.  Returns K or < 0. */
int aak_test_host_plan(const double *params, double T_fit, int length, double dt, int LISA,
                       int backint, int n_max, double *traj_out, double *knots_out, int cap,
                       int *info, double *dparams);

/* Device knot table of the host stage (y of e, v, M, S | their natural-spline
 * coefficients c | head knot times); grid[0..3] = table knots K, head offset,
 * head length, flat flag (1: the knots after the table bound flat intervals of
 * value y[K-1] and are not stored).  Returns the table length. */
int aak_test_knot_table(const double *params, double T_fit, int length, double dt, int LISA, int backint,
                        double *table, int cap, int *grid);

/* Knot interval and spline values (e, v, M, S) at n times, as the device
 * evaluates them: k_out[n], y_out[4n]. */
int aak_test_knot_eval(const double *params, double T_fit, int length, double dt, int LISA, int backint,
                       const double *t, int n, int *k_out, double *y_out);

/* J_{n-2..n+2}(n e) by the mode sum's series evaluation for eccentricities up
 * to e_bound; returns the series length used. */
int aak_test_bessel_five(int n, double e, double e_bound, double *J);

/* Mode-sum amplitude factors alpha_n, beta_n, gamma_n at e = e_ref + d from
 * their Taylor expansion about e_ref (order chosen for |d| <= del and
 * harmonics up to n_max), out[0..2], and from the direct Bessel evaluation,
 * out[3..5].  Returns the Taylor order. */
int aak_test_taylor_amplitudes(int n_max, int n, double e_ref, double del, double d, double *out);

/* One evaluation of the inspiral rates (dp/dt, de/dt, dcos(iota)/dt). */
int aak_test_inspiral_rates(double p, double e, double cosi, double a, double *out);

/* The inner-product stage alone on two given real series of length `length`
 * (the handle's data and weights from aak_input_data): like[0..1] =
 * (<h|d>, <h|h>) as aak_Likelihood returns them, spec (optional) =
 * [2][length/2+1] weighted complex spectra, *fused (optional) = 1 when the
 * fused FFT ran, 0 for the rocFFT path.  Invalidates the last waveform. */
struct aak_handle_s;
int aak_test_inner_product(struct aak_handle_s *h, const double *hI, const double *hII, double *like, double *spec,
                           int *fused);

/* on = 1: every phase fold of later waveforms on this handle skips the
 * parallel resolution and takes the serial replay (the fallback used when the
 * fold's bookkeeping overflows), so that path can be checked.  on = 2 (the
 * single-pass fold only): the middle block of every phase fails its own
 * check after its look-back, so the failure reaches the blocks after it and
 * the phase is replayed once every block has finished.  0: off. */
int aak_test_force_fold_replay(struct aak_handle_s *h, int on);

/* Form of the exact phase fold for later waveforms on this handle: 0 (the
 * default) one pass in which every block looks back over the blocks before
 * it for its exact start, 1 three passes (classification, a one-workgroup
 * chain over all blocks, write).  Both reproduce the sequential FP64 sum of
 * AAK.cc:244-264 bit for bit. */
int aak_test_fold_mode(struct aak_handle_s *h, int mode);

/* The Phidot guard (k_guard) alone on rate[0..n): every sample i in
 * (i0, guard_last] with rate <= 0 takes the rate of the last earlier sample
 * in (i0, i] that is not <= 0 (NaN counts as such), or of i0 (AAK.cc:254).
 * rate is rewritten in place; returns the number of replaced samples. */
int aak_test_phidot_guard(double *rate, int n, int i0, int guard_last);

/* Copies the dense arrays of the last aak_gpu_gen_AAK to the host:
 * out[10*(length+1)] = e | v | M | S | gimdot | Phidot | alpdot | gim | Phi | alp. */
int aak_test_get_dense(struct aak_handle_s *h, double *out);

#ifdef __cplusplus
}
#endif
#endif
