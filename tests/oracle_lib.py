"""ctypes bindings to the CPU oracle (oracle/liboracle.so) and to the
reference's own trajectory subset (oracle/_ref/libaakref.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, to check results or time the CPU baseline;
never by the product.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libaakref.so")

_d = C.c_double
_i = C.c_int
_pd = C.POINTER(C.c_double)
_pi = C.POINTER(C.c_int)


def _ptr(a):
    return a.ctypes.data_as(_pd)


class TrajData(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "t", "p", "ecc", "cosiota", "E", "Lz", "Q", "p3", "p4",
        "beta_wilkins", "zedminus", "betazedplus")]


_oracle = None
_ref = None


def oracle():
    """Load liboracle.so (built by `make -C oracle`)."""
    global _oracle
    if _oracle is None:
        lib = C.CDLL(ORACLE_SO)
        sig = {
            "o_iekg": (_i, [_d, _d, _d, _d, _i, _pd]),
            "o_param_derivs": (_i, [_d, _d, _d, _d, _pd]),
            "o_ckg": (_i, [_d, _d, _d, _pd]),
            "o_plso_table": (_i, [_d, _d, _pd]),
            "o_eccentric": (_i, [_d, _d, _d, _d, _d, _i, C.POINTER(TrajData), _pi]),
            "o_rf": (_d, [_d, _d, _d]), "o_rd": (_d, [_d, _d, _d]),
            "o_rj": (_d, [_d, _d, _d, _d]), "o_rc": (_d, [_d, _d]),
            "o_ellf": (_d, [_d, _d]), "o_elle": (_d, [_d, _d]),
            "o_ellpi": (_d, [_d, _d, _d]),
            "o_quartic_roots": (_i, [_pd, _pd, _pd]),
            "o_dvdt": (_d, [_d] * 6), "o_dedt": (_d, [_d] * 6),
            "o_dtdm": (_d, [_d] * 4), "o_drdm": (_d, [_d] * 4),
            "o_dthetadm": (_d, [_d] * 4), "o_dphidm": (_d, [_d] * 4),
            "o_omegaphi": (_d, [_d] * 5),
            "o_parmap": (_i, [_pd, _pd, _d, _d, _d, _d, _d]),
            "o_sol_fun": (None, [_pd, _pd, _d, _d, _d, _pd]),
            "o_spline_natural": (_i, [_pd, _pd, _i, _pd]),
            "o_spline_eval": (_d, [_pd, _pd, _pd, _i, _d]),
            "o_polyfit": (_i, [_pd, _pd, _pd, _i]),
            "o_bessel_jn_array": (_i, [_i, _i, _d, _pd]),
            "o_parang_phi0": (_d, [_d, _d]),
            "o_rotcoeff": (None, [_pd, _d, _d, _d, _d, _d, _d]),
            "o_sparse_stage": (_i, [_d, _i, _d, _d, _d, _d, _d, _d, _d, _d,
                                    _pd, _pd, _pd, _pd, _pd, _pd, _pi]),
            "o_pn_knots": (_i, [_d, _i, _pd, _pd, _pd, _d, _pd, _d, _pd, _pd, _i, _i,
                                _pd, _pd, _pd, _pd, _pd, _i, _pi]),
            "o_pn_j0": (_i, [_d, _i, _pd, _i]),
            "o_aak_wave": (_i, [_pd, _d, _i, _d, _i, _i, _i, _pd, _pd, _pd, _pd, _pd]),
            "o_waveform_from_dense": (_i, [_pd, _i, _d, _i, _d, _d, _d, _d, _d, _d, _d,
                                           _i, _i, _i, _pd, _pd, _pd]),
            "o_mode_abc": (None, [_i, _d, _d, _d, _pd]),
            "o_aak_phase": (_i, [_pd, _d, _i, _d, _i, _pd]),
            "o_last_error": (C.c_char_p, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _oracle = lib
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    """Load the compiled reference subset (oracle/_ref/libaakref.so)."""
    global _ref
    if _ref is None:
        lib = C.CDLL(REF_SO)
        sig = {
            "ref_eccentric": (_i, [_d, _d, _d, _d, _d, _i, _pd, _pi]),
            "ref_plso_table": (None, [_d, _d, _pd]),
            "ref_iekg": (None, [_d, _d, _d, _d, _pd]),
            "ref_param_derivs": (None, [_d, _d, _d, _d, _pd]),
            "ref_ckg": (None, [_d, _d, _d, _pd]),
            "ref_rf": (_d, [_d, _d, _d]), "ref_rd": (_d, [_d, _d, _d]),
            "ref_rj": (_d, [_d, _d, _d, _d]), "ref_rc": (_d, [_d, _d]),
            "ref_ellf": (_d, [_d, _d]), "ref_elle": (_d, [_d, _d]),
            "ref_ellpi": (_d, [_d, _d, _d]),
            "ref_quartic_roots": (None, [_pd, _pd, _pd]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _ref = lib
    return _ref


# ---------------------------------------------------------------- wrappers
PARAM_ORDER = ("iota", "s", "p", "e", "M", "mu", "gamma", "psi", "alph",
               "theta_S", "phi_S", "theta_K", "phi_K", "D")


def o_iekg(p, e, cosi, a, freq=True):
    out = np.zeros(16)
    rc = oracle().o_iekg(p, e, cosi, a, int(freq), _ptr(out))
    if rc:
        raise RuntimeError(oracle().o_last_error().decode())
    return out


def r_iekg(p, e, cosi, a):
    out = np.zeros(16)
    ref().ref_iekg(p, e, cosi, a, _ptr(out))
    return out


def o_eccentric(p, e, cosi, a, dtstart, maxsteps=100):
    traj = (TrajData * (maxsteps + 1))()
    steps = C.c_int(0)
    flag = oracle().o_eccentric(p, e, cosi, a, dtstart, maxsteps, traj, C.byref(steps))
    n = steps.value
    arr = np.array([[getattr(traj[i], f) for f, _ in TrajData._fields_]
                    for i in range(1, n + 1)]).reshape(n, 12)
    return flag, arr


def r_eccentric(p, e, cosi, a, dtstart, maxsteps=100):
    out = np.zeros(12 * maxsteps)
    steps = C.c_int(0)
    flag = ref().ref_eccentric(p, e, cosi, a, dtstart, maxsteps, _ptr(out), C.byref(steps))
    n = steps.value
    return flag, out[:12 * n].reshape(n, 12)


def o_aak_wave(pars, T_fit=1.0, length=2 ** 14, dt=10.0, LISA=True, backint=False,
               n_max=0, want_dense=False):
    """Full CPU AAK. pars: dict or sequence in PARAM_ORDER."""
    if isinstance(pars, dict):
        pars = [pars[k] for k in PARAM_ORDER]
    pv = np.ascontiguousarray(pars, dtype=np.float64)
    t = np.zeros(length)
    hI = np.zeros(length)
    hII = np.zeros(length)
    meta = np.zeros(8)
    dense = np.zeros(10 * (length + 1)) if want_dense else None
    rc = oracle().o_aak_wave(_ptr(pv), T_fit, length, dt, int(LISA), int(backint), int(n_max),
                             _ptr(t), _ptr(hI), _ptr(hII),
                             _ptr(dense) if want_dense else None, _ptr(meta))
    if rc != 0:
        raise RuntimeError("oracle AAK failed (%d): %s" % (rc, oracle().o_last_error().decode()))
    res = {"t": t, "hI": hI, "hII": hII, "steps": int(meta[0]), "nmodes": int(meta[1]),
           "i_plunge": int(meta[2]), "i_buffer": int(meta[3]), "K": int(meta[4]),
           "i0": int(meta[5]), "Phi0": meta[6], "zeta": meta[7]}
    if want_dense:
        d = dense.reshape(10, length + 1)
        for k, name in enumerate(("t", "e", "v", "M", "S", "gim", "Phi", "alp", "nu", "gimdot")):
            res["dense_" + name] = d[k]
        res["dense"] = dense
    return res


PHASE_NAMES = ("t", "phase_r", "phase_theta", "phase_phi", "omega_r", "omega_theta", "omega_phi", "eccentricity")


def o_aak_phase(pars, T_fit=1.0, length=2 ** 14, dt=10.0, backint=False):
    """AAK phases and frequencies (AAKwrapper.phase)."""
    if isinstance(pars, dict):
        pars = [pars[k] for k in PARAM_ORDER]
    pv = np.ascontiguousarray(pars, dtype=np.float64)
    out = np.zeros(8 * length)
    rc = oracle().o_aak_phase(_ptr(pv), T_fit, length, dt, int(backint), _ptr(out))
    if rc != 0:
        raise RuntimeError("oracle AAK phase failed (%d): %s" % (rc, oracle().o_last_error().decode()))
    return dict(zip(PHASE_NAMES, out.reshape(8, length)))


def o_knots(pars, T_fit=1.0, length=2 ** 14, dt=10.0, backint=False):
    """Sparse stage + PNevolution knots of the oracle (AAKpy.cc:16-44,
    AAK.cc:27-237): dict with the knot arrays, the sparse trajectory, the
    12-entry par vector and info = (i0, j0, j_min, j_max, j_max_temp,
    j_start, j_end, plunged, points)."""
    if isinstance(pars, dict):
        pars = [pars[k] for k in PARAM_ORDER]
    iota, s, p, e, M, mu, gamma, psi, alph, qS, phS, qK, phK, D = [float(x) for x in pars]
    sp = [np.zeros(100) for _ in range(5)]
    Phi0 = C.c_double(0)
    flag = C.c_int(0)
    steps = oracle().o_sparse_stage(T_fit, length, dt, iota, s, p, e, M, mu, psi, *[_ptr(a) for a in sp],
                                    C.byref(Phi0), C.byref(flag))
    if steps < 3:
        raise RuntimeError("oracle sparse stage failed (%d)" % steps)
    e_traj, v_map, M_map, s_map, dt_map = sp
    msun = 6.67259e-11 * 1.98892e30 / (299792458.0 * 299792458.0 * 299792458.0)
    par = np.array([mu * msun, M_map[0] * msun, s_map[0], e_traj[0], iota, gamma, Phi0.value,
                    qS, phS, qK, phK, alph])
    j0 = oracle().o_pn_j0(dt, length, _ptr(dt_map), steps)
    cap = 2 * j0 + 8
    kn = np.zeros(5 * cap)
    info = np.zeros(9, dtype=np.int32)
    K = oracle().o_pn_knots(dt, length, _ptr(par), _ptr(e_traj), _ptr(v_map), M, _ptr(M_map), s, _ptr(s_map),
                            _ptr(dt_map), steps, int(backint), *[_ptr(kn[k * cap:]) for k in range(5)], cap,
                            info.ctypes.data_as(_pi))
    if K < 0:
        raise RuntimeError("oracle knots failed (%d): %s" % (K, oracle().o_last_error().decode()))
    kn = kn.reshape(5, cap)[:, :K]
    return {"t": kn[0], "e": kn[1], "v": kn[2], "M": kn[3], "S": kn[4], "K": K, "steps": steps,
            "e_traj": e_traj[:steps], "v_map": v_map[:steps], "M_map": M_map[:steps], "s_map": s_map[:steps],
            "dt_map": dt_map[:steps], "par": par, "info": info, "M_phys": M, "S_phys": s}
