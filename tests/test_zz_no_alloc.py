"""aak_inner_product_noalloc: the likelihood's inner product on caller-owned
device buffers with no device allocation (scratch in a file-scope device
array).  Its output must equal the handle-based inner product
(aak_test_inner_product after aak_input_data, the same kernels) on two calls
in a row, the second at a smaller length, and lengths beyond its maximum
must be refused with a message.  Device buffers come from the HIP runtime
the library itself runs on (hipMalloc through ctypes), so one runtime serves
the whole process."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(L):
    f = L.lib().aak_inner_product_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p] * 6 + [C.c_int, C.c_void_p, C.c_void_p]
    return f


class _Hip:
    """hipMalloc/hipMemcpy/hipFree of the HIP runtime libaak_hip.so links."""

    def __init__(self):
        from emri_kludge_suite_amd import _lib
        _lib.lib()  # loads libamdhip64.so.7 beside the library
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, x):
        x = np.ascontiguousarray(x)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(x.nbytes, 8)) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, x.ctypes.data, x.nbytes, 1) == 0  # host to device
        return p.value

    def download(self, p, n):
        out = np.zeros(n)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def _case(rng, N):
    hI, hII = rng.normal(size=N), rng.normal(size=N)
    nb = N // 2 + 1
    w1, w2 = rng.uniform(0.5, 2.0, nb), rng.uniform(0.5, 2.0, nb)
    dI = (rng.normal(size=nb) + 1j * rng.normal(size=nb)) * w1
    dII = (rng.normal(size=nb) + 1j * rng.normal(size=nb)) * w2
    return hI, hII, dI, dII, w1, w2


def _reference(N, case):
    from emri_kludge_suite_amd import GPUAAK, _lib
    hI, hII, dI, dII, w1, w2 = case
    g = GPUAAK(1.0, N // 10, N, 100.0, 10.0, True, False)
    g.input_data(dI, dII, w1, w2)
    like = np.zeros(2)
    fused = C.c_int(0)
    _lib.check(_lib.lib().aak_test_inner_product(g._h, _lib.dptr(hI), _lib.dptr(hII), _lib.dptr(like), None,
                                                 C.byref(fused)), "inner product")
    assert fused.value == 1
    return like


def test_noalloc_inner_product_equals_the_handle_path():
    from emri_kludge_suite_amd import _lib
    f = _noalloc(_lib)
    hip = _Hip()
    rng = np.random.default_rng(61)
    try:
        for N in (2 ** 21, 2 ** 16):
            print("case noalloc N=%d" % N, flush=True)
            case = _case(rng, N)
            want = _reference(N, case)
            dev = [hip.upload(x) for x in case]
            out = hip.upload(np.full(2, np.nan))
            assert hip.h.hipDeviceSynchronize() == 0
            rc = f(*dev, N, out, None)
            assert rc == 0, _lib.last_error()
            got = hip.download(out, 2)
            print("  noalloc", got, "handle", want, "bitwise", bool(np.array_equal(got, want)), flush=True)
            np.testing.assert_allclose(got, want, rtol=1e-14, atol=0)
    finally:
        hip.free()


def test_noalloc_refuses_lengths_beyond_its_maximum():
    from emri_kludge_suite_amd import _lib
    f = _noalloc(_lib)
    hip = _Hip()
    try:
        p = hip.upload(np.zeros(16))
        for N in (2 ** 22, 3000, 128):
            assert f(p, p, p, p, p, p, N, p, None) == -2
            assert "power of two" in _lib.last_error()
    finally:
        hip.free()
