"""Python face of the GPU AAK path, mirroring the reference Cython class
``gpuAAK.GPUAAK`` (pygpuAAK/GPUAAK.pyx:50-93): same constructor arguments,
same methods, same return shapes.  Calls go through the C ABI of
include/gpuAAK.h; errors raise RuntimeError instead of terminating the process.
"""
import ctypes as C

import numpy as np

from . import _lib


def OmegaPhi(v, e, cosiota, s, M):
    """|Omega_phi| of the AK orbit (gpuAAK.h:13)."""
    return _lib.lib().aak_OmegaPhi(float(v), float(e), float(cosiota), float(s), float(M))


def _knot_outputs(call, cap):
    arrs = [np.zeros(cap) for _ in range(5)]
    ip, ib, K = C.c_int(0), C.c_int(0), C.c_int(0)
    tclip, dtl = C.c_double(0), C.c_double(0)
    rc = call([_lib.dptr(a) for a in arrs], C.byref(ip), C.byref(ib), C.byref(tclip), C.byref(dtl), C.byref(K))
    return rc, arrs, ip.value, ib.value, tclip.value, dtl.value, K.value


def PNevolution(timestep, vlength, par, e_traj, v_map, M_phys, M_map, S_phys, S_map, dt_map, steps, backint,
                cap=None):
    """AK evolution, fits and plunge search of the host stage (gpuAAK.h:15-18).

    Returns (t_in, e_in, v_in, M_in, S_in, i_plunge, i_buffer, t_clip,
    interp_timestep, init_length); the knot arrays hold init_length values.
    ``cap`` bounds the knot count (default: twice the knots the grid needs,
    as the reference sizes its evolution arrays)."""
    L = _lib.lib()
    ins = [np.ascontiguousarray(x, dtype=np.float64) for x in (par, e_traj, v_map, M_map, S_map, dt_map)]
    steps = int(steps)
    if cap is None:
        dt_large = ins[5][1]
        j0 = int((int(vlength) * float(timestep) - ins[5][steps - 1]) / dt_large + steps + 1)
        cap = 2 * max(j0, 1) + 8
    par_, e_, v_, M_, S_, d_ = [_lib.dptr(x) for x in ins]

    def call(out, ip, ib, tclip, dtl, K):
        return L.aak_PNevolution(*out, int(cap), float(timestep), int(vlength), par_, e_, v_, float(M_phys), M_,
                                 float(S_phys), S_, d_, steps, ip, ib, int(bool(backint)), tclip, dtl, K)

    rc, arrs, ip, ib, tclip, dtl, K = _knot_outputs(call, int(cap))
    _lib.check(rc, "PNevolution")
    return tuple(a[:K] for a in arrs) + (ip, ib, tclip, dtl, K)


class GPUAAK:
    """GPUAAK(T_fit, init_length, length, init_dt, dt, LISA, backint)

    Owns one device waveform slot of ``length`` samples, an rFFT plan and the
    weighted data of both TDI-like channels (GPUAAK.pyx:56-59).
    """

    def __init__(self, T_fit, init_length, length, init_dt, dt, LISA, backint):
        self.length = int(length)
        self._h = C.c_void_p()
        L = _lib.lib()
        _lib.check(L.aak_create(float(T_fit), int(init_length), self.length, float(init_dt),
                                float(dt), int(bool(LISA)), int(bool(backint)), C.byref(self._h)),
                   "GPUAAK()")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.lib().aak_destroy(h)
            except Exception:
                pass
            self._h = None

    def close(self):
        self.__del__()

    # ------------------------------------------------------------------ API
    def input_data(self, hI_f_, hII_f_, channel_ASDinv1_, channel_ASDinv2_):
        """Weighted data and inverse ASDs of both channels (GPUAAK.pyx:61-67)."""
        a = np.ascontiguousarray(hI_f_, dtype=np.complex128)
        b = np.ascontiguousarray(hII_f_, dtype=np.complex128)
        w1 = np.ascontiguousarray(channel_ASDinv1_, dtype=np.float64)
        w2 = np.ascontiguousarray(channel_ASDinv2_, dtype=np.float64)
        n = len(a)
        if not (len(b) == n and len(w1) == n and len(w2) == n):
            raise ValueError("input_data: all four arrays need the same length")
        _lib.check(_lib.lib().aak_input_data(self._h, a.view(np.float64).ctypes.data_as(_lib._pd),
                                             b.view(np.float64).ctypes.data_as(_lib._pd),
                                             _lib.dptr(w1), _lib.dptr(w2), n), "input_data")

    def run_phase_trajectory(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D):
        """Host stage alone: sparse inspiral, frequency map and PNevolution
        (AAK_manager.hh:118-132); the knots stay in the handle."""
        _lib.check(_lib.lib().aak_run_phase_trajectory(self._h, *[float(x) for x in (
            iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D)]), "run_phase_trajectory")

    def GetTrajectory(self, cap=1 << 22):
        """Knots and PNevolution outputs kept by the last run_phase_trajectory
        or gpu_gen_AAK: (t_in, e_in, v_in, M_in, S_in, i_plunge, i_buffer,
        t_clip, interp_timestep, init_length)."""
        L = _lib.lib()

        def call(out, ip, ib, tclip, dtl, K):
            return L.aak_get_trajectory(self._h, *out, int(cap), ip, ib, tclip, dtl, K)

        rc, arrs, ip, ib, tclip, dtl, K = _knot_outputs(call, int(cap))
        _lib.check(rc, "GetTrajectory")
        return tuple(a[:K] for a in arrs) + (ip, ib, tclip, dtl, K)

    def gpu_gen_AAK(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D):
        """Generate the waveform on the device (GPUAAK.pyx:69-74)."""
        _lib.check(_lib.lib().aak_gpu_gen_AAK(self._h, *(float(x) for x in (
            iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D))), "gpu_gen_AAK")

    def GetWaveform(self, is_Fourier=False):
        """(t, hI, hII), each length+2 doubles (GPUAAK.pyx:76-83)."""
        n = self.length + 2
        t = np.zeros(n)
        hI = np.zeros(n)
        hII = np.zeros(n)
        _lib.check(_lib.lib().aak_get_waveform(self._h, _lib.dptr(t), _lib.dptr(hI), _lib.dptr(hII)),
                   "GetWaveform")
        return (t, hI, hII)

    def Likelihood(self):
        """[4<h|d>, 4<h|h>] of the last waveform (GPUAAK.pyx:85-88)."""
        like = np.zeros(2)
        _lib.check(_lib.lib().aak_likelihood(self._h, _lib.dptr(like)), "Likelihood")
        return like

    def WaveformThroughLikelihood(self, iota, s, p, e, M, mu, gamma, psi, alph, theta_S,
                                  phi_S, theta_K, phi_K, D, return_waveform=False):
        """gpu_gen_AAK then Likelihood (GPUAAK.pyx:90-93)."""
        self.gpu_gen_AAK(iota, s, p, e, M, mu, gamma, psi, alph, theta_S, phi_S, theta_K, phi_K, D)
        if return_waveform:
            return (0.0, 0.0)
        return self.Likelihood()

    # ------------------------------------------------------------ extensions
    def LikelihoodBatch(self, params, return_status=False, nmax=None):
        """Likelihoods of many templates against the current data.

        ``params``: (n, 14) array in the gpu_gen_AAK order.  Returns an (n, 2)
        array whose rows equal WaveformThroughLikelihood(*params[i]); rows of
        templates whose host stage fails are NaN (their negative status codes
        are returned as well with ``return_status=True``).  The host stages run
        on a thread pool while the GPU works on earlier templates.  ``nmax``
        (optional, n ints): a harmonic count per template (0: the reference's
        rule), instead of the handle's setting.
        """
        P = np.ascontiguousarray(params, dtype=np.float64)
        if P.ndim != 2 or P.shape[1] != 14:
            raise ValueError("LikelihoodBatch: params must have shape (n, 14)")
        n = P.shape[0]
        out = np.zeros((n, 2))
        st = np.zeros(n, dtype=np.int32)
        if nmax is None:
            rc = _lib.lib().aak_likelihood_batch(self._h, _lib.dptr(P), n, _lib.dptr(out), _lib.iptr(st))
        else:
            nm = np.ascontiguousarray(nmax, dtype=np.int32)
            if nm.shape != (n,):
                raise ValueError("LikelihoodBatch: nmax needs one entry per template")
            rc = _lib.lib().aak_likelihood_batch_nmax(self._h, _lib.dptr(P), _lib.iptr(nm), n, _lib.dptr(out),
                                                      _lib.iptr(st))
        _lib.check(rc, "LikelihoodBatch")
        return (out, st) if return_status else out

    def set_precision(self, mode):
        """Mode-sum arithmetic: "fp64" (default, parity with the reference) or
        "mixed" (FP32 harmonic loop, mismatch below 1e-6)."""
        m = {"fp64": 0, "mixed": 1}.get(mode, mode)
        _lib.check(_lib.lib().aak_set_precision(self._h, int(m)), "set_precision")

    def set_host_threads(self, n):
        """Threads of the batch host pool (0: AAK_HOST_THREADS or the core count, at most 16)."""
        _lib.check(_lib.lib().aak_set_host_threads(self._h, int(n)), "set_host_threads")

    def set_lanes(self, n):
        """Concurrent streams of LikelihoodBatch (1..8); results do not depend on it."""
        _lib.check(_lib.lib().aak_set_lanes(self._h, int(n)), "set_lanes")

    def set_tuning(self, feed_depth=-1, exp_shift=-1):
        """Batch pipeline knobs: short host plans kept ready for the GPU
        (1..256), mode-sum expansion blocks of 256 << exp_shift samples (0..3;
        -2 restores the precision's default); < 0 (-1) keeps a knob."""
        _lib.check(_lib.lib().aak_set_tuning(self._h, int(feed_depth), int(exp_shift)), "set_tuning")

    def set_map_init(self, mode, model=None):
        """Start point of the frequency map: "solver" (the reference's) or
        "learned" (the network `model` plus a Newton polish, solver fallback)."""
        m = {"solver": 0, "learned": 1}[mode] if isinstance(mode, str) else int(mode)
        self._map_model = model  # the handle uses it; keep it alive
        _lib.check(_lib.lib().aak_set_map_init(self._h, m, model._m if model is not None else None), "set_map_init")

    def set_nmax(self, n_max):
        """Override the harmonic count (0 restores (int)(30 e0), min 4)."""
        _lib.check(_lib.lib().aak_set_nmax(self._h, int(n_max)), "set_nmax")

    def info(self):
        """Diagnostics of the last gpu_gen_AAK."""
        buf = np.zeros(12, dtype=np.int32)
        _lib.lib().aak_get_info(self._h, _lib.iptr(buf), 12)
        keys = ("steps", "nmodes", "i_plunge", "i_buffer", "K", "i0", "fold_fallbacks",
                "nonpositive_phidot", "j0", "plunged", "per_sample_blocks", "split_blocks")
        return dict(zip(keys, (int(x) for x in buf)))

    def set_profiling(self, on=True):
        """Record HIP events around each stage on the library's stream."""
        _lib.check(_lib.lib().aak_set_profiling(self._h, int(bool(on))), "set_profiling")

    def timings(self):
        """Stage times (ms) of the last gpu_gen_AAK / Likelihood.

        fft / dot: for power-of-two lengths the column pass of the fused FFT,
        then the row pass with the weighting and both sums; otherwise the
        rocFFT transform, then the weighted reduction.
        """
        buf = np.zeros(6)
        _lib.lib().aak_get_timings(self._h, _lib.dptr(buf), 6)
        keys = ("host", "rates", "fold", "modesum", "fft", "dot")
        return dict(zip(keys, (float(x) for x in buf)))

    def batch_timings(self):
        """Stage time sums (ms) over the last LikelihoodBatch run with profiling on."""
        buf = np.zeros(5)
        _lib.check(_lib.lib().aak_get_batch_timings(self._h, _lib.dptr(buf), 5), "batch_timings")
        keys = ("rates", "fold", "modesum", "fft", "dot")
        return dict(zip(keys, (float(x) for x in buf)))

    def batch_host_ms(self):
        """Host-stage time (ms) of the last LikelihoodBatch: (sum over templates, longest)."""
        a, b = np.zeros(1), np.zeros(1)
        _lib.check(_lib.lib().aak_get_batch_host_ms(self._h, _lib.dptr(a), _lib.dptr(b)), "batch_host_ms")
        return float(a[0]), float(b[0])

    def pool_stats(self):
        """Host pool of the last LikelihoodBatch: threads, host ms summed over
        the templates, longest template's host ms, wall ms of the call, busy
        fraction, and the issuing thread's ms enqueueing and waiting for plans."""
        buf = np.zeros(7)
        _lib.check(_lib.lib().aak_get_pool_stats(self._h, _lib.dptr(buf), 7), "pool_stats")
        return dict(zip(("threads", "host_ms_sum", "host_ms_max", "wall_ms", "busy", "issue_ms", "wait_ms"),
                        (float(x) for x in buf)))

    def inner_product_stage(self, hI, hII):
        """Runs the inner-product stage alone on two real series (test hook).

        Returns ((<h|d>, <h|h>), weighted spectra [2][length/2+1] complex,
        True if the fused FFT ran).  Invalidates the last waveform.
        """
        n = self.length
        a = np.ascontiguousarray(hI, dtype=np.float64)
        b = np.ascontiguousarray(hII, dtype=np.float64)
        if a.shape != (n,) or b.shape != (n,):
            raise ValueError("hI and hII must have the handle's length")
        like = np.zeros(2)
        spec = np.zeros(4 * (n // 2 + 1))
        fused = np.zeros(1, dtype=np.int32)
        _lib.check(_lib.lib().aak_test_inner_product(self._h, _lib.dptr(a), _lib.dptr(b), _lib.dptr(like),
                                                     _lib.dptr(spec), _lib.iptr(fused)), "inner_product_stage")
        spec = spec[0::2] + 1j * spec[1::2]
        return (float(like[0]), float(like[1])), spec.reshape(2, n // 2 + 1), bool(fused[0])

    def force_fold_replay(self, on=True):
        """Route every later phase fold through the serial replay (test hook);
        on=2: the middle block of each phase fails after its look-back."""
        mode = 2 if on == 2 else int(bool(on))
        _lib.check(_lib.lib().aak_test_force_fold_replay(self._h, mode), "force_fold_replay")

    def fold_mode(self, mode):
        """0: single-pass look-back fold (default); 1: three-pass fold (test hook)."""
        _lib.check(_lib.lib().aak_test_fold_mode(self._h, int(mode)), "fold_mode")

    def dense_arrays(self):
        """Device dense arrays of the last waveform (test hook)."""
        L = self.length + 1
        out = np.zeros(10 * L)
        _lib.check(_lib.lib().aak_test_get_dense(self._h, _lib.dptr(out)), "dense_arrays")
        out = out.reshape(10, L)
        names = ("e", "v", "M", "S", "gimdot", "Phidot", "alpdot", "gim", "Phi", "alp")
        return dict(zip(names, out))
