# cython: language_level=3, boundscheck=False, wraparound=False
"""gpuAAK -- the reference's Cython module (pygpuAAK/GPUAAK.pyx:50-93) on the
C ABI of libaak_hip.so (include/gpuAAK.h).

Class and method signatures are the reference's, so its caller keeps
``from gpuAAK import GPUAAK`` (pygpuAAK/pygpuAAK.py:5).  Status codes of the
library become RuntimeError instead of the reference's exit().  Beyond the
reference: LikelihoodBatch (many templates per call), run_phase_trajectory /
GetTrajectory, and the free functions OmegaPhi and PNevolution of the
reference's gpuAAK.h.
"""
import numpy as np

cdef extern from "gpuAAK.h":
    ctypedef struct aak_handle_s
    ctypedef aak_handle_s* aak_handle
    int aak_create(double T_fit, int init_length, int length, double init_dt, double dt,
                   int LISA, int backint, aak_handle* out)
    int aak_destroy(aak_handle h)
    int aak_input_data(aak_handle h, const double* hI_f, const double* hII_f,
                       const double* ASDinv1, const double* ASDinv2, int len)
    int aak_gpu_gen_AAK(aak_handle h, double iota, double s, double p, double e, double M,
                        double mu, double gamma, double psi, double alph, double theta_S,
                        double phi_S, double theta_K, double phi_K, double D)
    int aak_likelihood(aak_handle h, double* like_out)
    int aak_get_waveform(aak_handle h, double* t, double* hI, double* hII)
    int aak_likelihood_batch(aak_handle h, const double* params, int n, double* like_out, int* status)
    int aak_run_phase_trajectory(aak_handle h, double iota, double s, double p, double e, double M,
                                 double mu, double gamma, double psi, double alph, double theta_S,
                                 double phi_S, double theta_K, double phi_K, double D)
    int aak_get_trajectory(aak_handle h, double* t_in, double* e_in, double* v_in, double* M_in,
                           double* S_in, int cap, int* i_plunge, int* i_buffer, double* t_clip,
                           double* interp_timestep, int* init_length)
    double aak_OmegaPhi(double v, double e, double cosiota, double s, double M)
    int aak_PNevolution(double* t_in, double* e_in, double* v_in, double* M_in, double* S_in, int cap,
                        double timestep, int vlength, const double* par, const double* e_traj,
                        const double* v_map, double M_phys, const double* M_map, double S_phys,
                        const double* S_map, const double* dt_map, int steps, int* i_plunge,
                        int* i_buffer, int backint, double* t_clip, double* interp_timestep,
                        int* init_length)
    const char* aak_last_error()


cdef int _check(int rc) except -1:
    if rc < 0:
        raise RuntimeError(aak_last_error().decode(errors="replace"))
    return rc


def OmegaPhi(double v, double e, double cosiota, double s, double M):
    """OmegaPhi (gpuAAK.h:13)."""
    return aak_OmegaPhi(v, e, cosiota, s, M)


def PNevolution(double timestep, int vlength, par, e_traj, v_map, double M_phys, M_map, double S_phys,
                S_map, dt_map, int steps, backint, int cap):
    """PNevolution (gpuAAK.h:15-18): returns (t_in, e_in, v_in, M_in, S_in,
    i_plunge, i_buffer, t_clip, interp_timestep, init_length); at most cap knots."""
    cdef double[::1] p_ = np.ascontiguousarray(par, dtype=np.float64)
    cdef double[::1] et = np.ascontiguousarray(e_traj, dtype=np.float64)
    cdef double[::1] vm = np.ascontiguousarray(v_map, dtype=np.float64)
    cdef double[::1] Mm = np.ascontiguousarray(M_map, dtype=np.float64)
    cdef double[::1] Sm = np.ascontiguousarray(S_map, dtype=np.float64)
    cdef double[::1] dm = np.ascontiguousarray(dt_map, dtype=np.float64)
    out = np.zeros((5, cap))
    cdef double[:, ::1] o = out
    cdef int ip = 0, ib = 0, K = 0
    cdef double tclip = 0, dtl = 0
    _check(aak_PNevolution(&o[0, 0], &o[1, 0], &o[2, 0], &o[3, 0], &o[4, 0], cap, timestep, vlength,
                           &p_[0], &et[0], &vm[0], M_phys, &Mm[0], S_phys, &Sm[0], &dm[0], steps,
                           &ip, &ib, 1 if backint else 0, &tclip, &dtl, &K))
    return tuple(out[a, :K].copy() for a in range(5)) + (ip, ib, tclip, dtl, K)


cdef class GPUAAK:
    cdef int length
    cdef aak_handle g

    def __cinit__(self, T_fit, init_length, length, init_dt, dt, LISA, backint):
        self.length = length
        self.g = NULL
        _check(aak_create(T_fit, init_length, length, init_dt, dt, 1 if LISA else 0, 1 if backint else 0,
                          &self.g))

    def __dealloc__(self):
        if self.g != NULL:
            aak_destroy(self.g)
            self.g = NULL

    def input_data(self, hI_f_, hII_f_, channel_ASDinv1_, channel_ASDinv2_):
        cdef double complex[::1] a = np.ascontiguousarray(hI_f_, dtype=np.complex128)
        cdef double complex[::1] b = np.ascontiguousarray(hII_f_, dtype=np.complex128)
        cdef double[::1] w1 = np.ascontiguousarray(channel_ASDinv1_, dtype=np.float64)
        cdef double[::1] w2 = np.ascontiguousarray(channel_ASDinv2_, dtype=np.float64)
        if not (b.shape[0] == a.shape[0] and w1.shape[0] == a.shape[0] and w2.shape[0] == a.shape[0]):
            raise ValueError("input_data: all four arrays need the same length")
        _check(aak_input_data(self.g, <double*>&a[0], <double*>&b[0], &w1[0], &w2[0], a.shape[0]))

    def gpu_gen_AAK(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D):
        _check(aak_gpu_gen_AAK(self.g, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S,
                               theta_K, phi_K, D))

    def run_phase_trajectory(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K,
                             phi_K, D):
        _check(aak_run_phase_trajectory(self.g, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S,
                                        theta_K, phi_K, D))

    def GetTrajectory(self, int cap=1 << 22):
        out = np.zeros((5, cap))
        cdef double[:, ::1] o = out
        cdef int ip = 0, ib = 0, K = 0
        cdef double tclip = 0, dtl = 0
        _check(aak_get_trajectory(self.g, &o[0, 0], &o[1, 0], &o[2, 0], &o[3, 0], &o[4, 0], cap, &ip, &ib,
                                  &tclip, &dtl, &K))
        return tuple(out[a, :K].copy() for a in range(5)) + (ip, ib, tclip, dtl, K)

    def GetWaveform(self, is_Fourier=False):
        t_ = np.zeros(self.length + 2)
        hI_ = np.zeros(self.length + 2)
        hII_ = np.zeros(self.length + 2)
        cdef double[::1] a = t_
        cdef double[::1] b = hI_
        cdef double[::1] c = hII_
        _check(aak_get_waveform(self.g, &a[0], &b[0], &c[0]))
        return (t_, hI_, hII_)

    def Likelihood(self):
        like_ = np.zeros(2)
        cdef double[::1] l = like_
        _check(aak_likelihood(self.g, &l[0]))
        return like_

    def WaveformThroughLikelihood(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K,
                                  phi_K, D, return_waveform=False):
        self.gpu_gen_AAK(iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D)
        if return_waveform:
            return (0.0, 0.0)
        return self.Likelihood()

    def LikelihoodBatch(self, params):
        """Rows of the 14 gpu_gen_AAK parameters -> rows [4<h|d>, 4<h|h>]
        (NaN where a template's host stage fails)."""
        cdef double[:, ::1] P = np.ascontiguousarray(np.atleast_2d(params), dtype=np.float64)
        if P.shape[1] != 14:
            raise ValueError("LikelihoodBatch: rows of 14 parameters")
        n = P.shape[0]
        like_ = np.zeros((n, 2))
        status = np.zeros(n, dtype=np.intc)
        if n == 0:
            return like_
        cdef double[:, ::1] l = like_
        cdef int[::1] st = status
        _check(aak_likelihood_batch(self.g, &P[0, 0], n, &l[0, 0], &st[0]))
        return like_
