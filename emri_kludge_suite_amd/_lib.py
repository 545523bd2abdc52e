"""Loader for the in-tree HIP library libaak_hip.so (C ABI of include/gpuAAK.h).

The library is required: there is no CPU fallback for any product entry point.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaak_hip.so")

_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)
_pf = C.POINTER(C.c_float)

_lib = None


class LibraryMissing(RuntimeError):
    pass


class SetPar(C.Structure):
    """aak_setpar: the settings/parameters of AAK_Waveform (KSTools.h:21-51)."""
    _fields_ = ([("path", C.c_char * 1024)] +
                [(n, C.c_int) for n in ("backint", "LISA", "traj", "SNR", "timing", "length")] +
                [(n, C.c_double) for n in ("dt", "p", "T", "f", "T_fit", "mu", "M", "s", "e", "iota", "gamma",
                                           "psi", "theta_S", "phi_S", "theta_K", "phi_K", "alpha", "D")])


def lib():
    """Return the loaded libaak_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            "libaak_hip.so not found at %s: build it with "
            "`make -C emri_kludge_suite_amd/csrc` (or __graft_entry__.build())" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    sig = {
        "aak_create": (_i, [_d, _i, _i, _d, _d, _i, _i, C.POINTER(C.c_void_p)]),
        "aak_destroy": (_i, [C.c_void_p]),
        "aak_input_data": (_i, [C.c_void_p, _pd, _pd, _pd, _pd, _i]),
        "aak_gpu_gen_AAK": (_i, [C.c_void_p] + [_d] * 14),
        "aak_likelihood": (_i, [C.c_void_p, _pd]),
        "aak_get_waveform": (_i, [C.c_void_p, _pd, _pd, _pd]),
        "aak_OmegaPhi": (_d, [_d] * 5),
        "aak_set_nmax": (_i, [C.c_void_p, _i]),
        "aak_get_info": (_i, [C.c_void_p, _pi, _i]),
        "aak_last_error": (C.c_char_p, []),
        "aak_test_host_plan": (_i, [_pd, _d, _i, _d, _i, _i, _i, _pd, _pd, _i, _pi, _pd]),
        "aak_test_knot_table": (_i, [_pd, _d, _i, _d, _i, _i, _pd, _i, _pi]),
        "aak_test_knot_eval": (_i, [_pd, _d, _i, _d, _i, _i, _pd, _i, _pi, _pd]),
        "aak_test_bessel_five": (_i, [_i, _d, _d, _pd]),
        "aak_test_taylor_amplitudes": (_i, [_i, _i, _d, _d, _d, _pd]),
        "aak_likelihood_batch": (_i, [C.c_void_p, _pd, _i, _pd, _pi]),
        "aak_likelihood_batch_nmax": (_i, [C.c_void_p, _pd, _pi, _i, _pd, _pi]),
        "aak_set_host_threads": (_i, [C.c_void_p, _i]),
        "aak_set_lanes": (_i, [C.c_void_p, _i]),
        "aak_set_tuning": (_i, [C.c_void_p, _i, _i]),
        "aak_set_precision": (_i, [C.c_void_p, _i]),
        "aak_gen_phases": (_i, [C.c_void_p, _pd] + [_pd] * 8),
        "aak_get_batch_timings": (_i, [C.c_void_p, _pd, _i]),
        "aak_test_inspiral_rates": (_i, [_d, _d, _d, _d, _pd]),
        "aak_test_get_dense": (_i, [C.c_void_p, _pd]),
        "aak_test_inner_product": (_i, [C.c_void_p, _pd, _pd, _pd, _pd, _pi]),
        "aak_test_force_fold_replay": (_i, [C.c_void_p, _i]),
        "aak_test_fold_mode": (_i, [C.c_void_p, _i]),
        "aak_test_phidot_guard": (_i, [_pd, _i, _i, _i]),
        "aak_get_batch_host_ms": (_i, [C.c_void_p, _pd, _pd]),
        "aak_set_profiling": (_i, [C.c_void_p, _i]),
        "aak_get_timings": (_i, [C.c_void_p, _pd, _i]),
        "aak_get_pool_stats": (_i, [C.c_void_p, _pd, _i]),
        "aak_PNevolution": (_i, [_pd] * 5 + [_i, _d, _i, _pd, _pd, _pd, _d, _pd, _d, _pd, _pd, _i, _pi, _pi, _i,
                                             _pd, _pd, _pi]),
        "aak_run_phase_trajectory": (_i, [C.c_void_p] + [_d] * 14),
        "aak_get_trajectory": (_i, [C.c_void_p] + [_pd] * 5 + [_i, _pi, _pi, _pd, _pd, _pi]),
        "aak_mlp_create": (_i, [_i, C.POINTER(C.c_void_p)]),
        "aak_mlp_destroy": (_i, [C.c_void_p]),
        "aak_mlp_set_params": (_i, [C.c_void_p, _pf, _pf]),
        "aak_mlp_set_norm": (_i, [C.c_void_p, _pf]),
        "aak_mlp_get_params": (_i, [C.c_void_p, _pf, _pf]),
        "aak_mlp_forward": (_i, [C.c_void_p, _pf, _i, _pf]),
        "aak_mlp_gradient": (_i, [C.c_void_p, _pf, _pf, _i, _pf, _pd]),
        "aak_mlp_train": (_i, [C.c_void_p, _pf, _pf, _i, _i, _i, C.c_float, C.c_float, _i, _pd]),
        "aak_mlp_predict_maps": (_i, [C.c_void_p, _pd, _i, _pd]),
        "aak_freqmap_batch": (_i, [_pd, _i, _i, C.c_void_p, _d, _i, _pd, _pi, _pd, _pd]),
        "aak_p_lso": (_d, [_d, _d, _d]),
        "aak_p_lso_batch": (_i, [_pd, _i, _pd, _i]),
        "aak_set_map_init": (_i, [C.c_void_p, _i, C.c_void_p]),
        "aak_get_inspiral": (_i, [C.c_void_p, _pd, _pd, _pd, _i]),
        "aak_kerr_constants": (_i, [_d, _d, _d, _d, _pd]),
        "aak_load_setpar": (_i, [C.c_char_p, C.POINTER(SetPar)]),
        "aak_elisa_noise": (_d, [_d]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def last_error():
    return lib().aak_last_error().decode(errors="replace")


def check(rc, what):
    if rc < 0:
        raise RuntimeError("%s failed (%d): %s" % (what, rc, last_error()))
    return rc


def dptr(a):
    return a.ctypes.data_as(_pd)


def iptr(a):
    return a.ctypes.data_as(_pi)


def fptr(a):
    return a.ctypes.data_as(_pf)
