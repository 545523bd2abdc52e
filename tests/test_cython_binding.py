"""The reference's Cython module rebuilt on the C ABI
(emri_kludge_suite_amd/cython/gpuAAK.pyx; reference pygpuAAK/GPUAAK.pyx:50-93):
it builds, imports under the reference's module name, and its host-only
functions agree with the ctypes mirror.  The GPU half is in
test_gpu_cython.py."""
import importlib.util
import os
import sys

import numpy as np
import pytest

import oracle_lib as ol

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CYDIR = os.path.join(ROOT, "emri_kludge_suite_amd", "cython")


def load_gpuAAK():
    """The compiled module, imported the way the reference's caller does
    (``from gpuAAK import GPUAAK``); built in place when absent."""
    import glob
    if not glob.glob(os.path.join(CYDIR, "gpuAAK*.so")):
        sys.path.insert(0, CYDIR)
        import build_gpuAAK
        build_gpuAAK.build()
    if CYDIR not in sys.path:
        sys.path.insert(0, CYDIR)
    import gpuAAK
    return gpuAAK


def test_module_exposes_the_reference_class():
    m = load_gpuAAK()
    for name in ("input_data", "gpu_gen_AAK", "GetWaveform", "Likelihood", "WaveformThroughLikelihood",
                 "LikelihoodBatch", "run_phase_trajectory", "GetTrajectory"):
        assert hasattr(m.GPUAAK, name)


def test_free_functions_match_the_ctypes_mirror():
    m = load_gpuAAK()
    from emri_kludge_suite_amd import OmegaPhi, PNevolution
    assert m.OmegaPhi(0.3, 0.2, 0.6, 0.5, 1.0) == OmegaPhi(0.3, 0.2, 0.6, 0.5, 1.0)
    pars = [0.5, 0.5, 8.0, 0.1, 1e6, 10.0, 1.0, 2.0, 3.0, 1.1, 4.0, 0.7, 5.0, 1.0]
    N, dt = 2 ** 18, 120.0
    r = ol.o_knots(pars, 1.0, N, dt)
    args = (dt, N, r["par"], r["e_traj"], r["v_map"], r["M_phys"], r["M_map"], r["S_phys"], r["s_map"],
            r["dt_map"], r["steps"], False)
    a = m.PNevolution(*args, 2 * r["K"])
    b = PNevolution(*args)
    for x, y in zip(a[:5], b[:5]):
        np.testing.assert_array_equal(x, y)
    assert a[5:] == b[5:]
    with pytest.raises(RuntimeError, match="knot buffer too small"):
        m.PNevolution(*args, 10)


def test_constructor_error_is_an_exception_without_a_gpu():
    m = load_gpuAAK()
    from emri_kludge_suite_amd import _lib
    import ctypes as C
    h = C.c_void_p()
    if _lib.lib().aak_create(1.0, 100, 1024, 100.0, 10.0, 1, 0, C.byref(h)) == 0:
        _lib.lib().aak_destroy(h)
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError, match="HIP device"):
        m.GPUAAK(1.0, 100, 1024, 100.0, 10.0, True, False)
