"""GPU versions of the reference's CPU AAK entry points wave() and phase()
(AAKwrapper/AAKwrapper.pyx:60-133, AAKpy.cc:16-88), same ``pars`` dictionary
and return tuples.  Each call runs on a cached device handle for its
(T_fit, length, dt, LISA, backint).  ``timing`` is the wall time in seconds,
as the reference reports its CPU time.
"""
import time

import numpy as np

from . import _lib
from .gpuAAK import GPUAAK

_ORDER = ("iota", "s", "p", "e", "M", "mu", "gamma", "psi", "alpha", "theta_S", "phi_S", "theta_K", "phi_K", "D")
_handles = {}


def _handle(pars):
    key = (float(pars["T_fit"]), int(pars["length"]), float(pars["dt"]), bool(pars["LISA"]), bool(pars["backint"]))
    g = _handles.get(key)
    if g is None:
        g = GPUAAK(key[0], max(1, key[1] // 100), key[1], 100.0, key[2], key[3], key[4])
        _handles[key] = g
    return g


def _params(pars):
    return np.array([float(pars[k]) for k in _ORDER])


def wave(pars={}):
    """(t, hI, hII, timing): the AAK waveform (AAKwrapper.pyx:60-88)."""
    t0 = time.perf_counter()
    g = _handle(pars)
    g.set_nmax(int(pars.get("n_max", 0)))
    g.gpu_gen_AAK(*_params(pars))
    t, hI, hII = g.GetWaveform()
    n = int(pars["length"])
    return t[:n].copy(), hI[:n].copy(), hII[:n].copy(), time.perf_counter() - t0


def phase(pars={}):
    """(t, phase_r, phase_theta, phase_phi, omega_r, omega_theta, omega_phi,
    eccentricity, timing): the AAK phases and frequencies (AAKwrapper.pyx:90-133)."""
    t0 = time.perf_counter()
    g = _handle(pars)
    n = int(pars["length"])
    out = [np.zeros(n) for _ in range(8)]
    P = _params(pars)
    _lib.check(_lib.lib().aak_gen_phases(g._h, _lib.dptr(P), *(_lib.dptr(a) for a in out)), "phase")
    return tuple(out) + (time.perf_counter() - t0,)
