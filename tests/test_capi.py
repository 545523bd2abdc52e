"""The C ABI of libaak_hip.so: every entry point that include/*.h declares is
exported, argument errors come back as status codes, and the product refuses
to run without a GPU instead of falling back to a CPU path.  No kernel is
launched here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(ROOT, "include", f) for f in sorted(os.listdir(os.path.join(ROOT, "include")))
           if f.endswith(".h")]


def _declared():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names.update(re.findall(r"\b(aak_\w+)\s*\(", text))
    return sorted(names)


@pytest.fixture(scope="module")
def L():
    from emri_kludge_suite_amd import _lib
    return _lib


def test_headers_declare_entry_points():
    names = _declared()
    for required in ("aak_create", "aak_destroy", "aak_input_data", "aak_gpu_gen_AAK", "aak_likelihood",
                     "aak_get_waveform", "aak_OmegaPhi", "aak_last_error"):
        assert required in names


def test_library_exports_every_declared_symbol(L):
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_null_handle_is_an_error(L):
    lib = L.lib()
    out = np.zeros(2)
    assert lib.aak_likelihood(None, L.dptr(out)) < 0
    assert lib.aak_gpu_gen_AAK(None, *([1.0] * 14)) < 0
    assert lib.aak_destroy(None) < 0
    assert "handle" in L.last_error()


def _gpu_visible(L):
    h = C.c_void_p()
    rc = L.lib().aak_create(1.0, 100, 1024, 100.0, 10.0, 1, 0, C.byref(h))
    if rc == 0:
        L.lib().aak_destroy(h)
        return True
    return False


def test_no_cpu_fallback_without_a_gpu(L):
    if _gpu_visible(L):
        pytest.skip("a GPU is visible: the no-device path cannot be exercised")
    from emri_kludge_suite_amd import GPUAAK
    with pytest.raises(RuntimeError, match="HIP device"):
        GPUAAK(1.0, 100, 1024, 100.0, 10.0, True, False)


def test_missing_library_fails_loudly(L, monkeypatch, tmp_path):
    monkeypatch.setattr(L, "_lib", None)
    monkeypatch.setattr(L, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(L.LibraryMissing):
        L.lib()
