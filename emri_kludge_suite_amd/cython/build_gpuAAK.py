"""Builds the Cython module gpuAAK (gpuAAK.pyx) in place, against the C ABI
header include/gpuAAK.h and the in-tree libaak_hip.so (found at run time
through an $ORIGIN-relative rpath).  Run: python emri_kludge_suite_amd/cython/build_gpuAAK.py"""
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)


def build():
    from Cython.Build import cythonize  # noqa: F401  (fails loudly when Cython is absent)
    from Cython.Compiler import Main
    pyx = os.path.join(HERE, "gpuAAK.pyx")
    c_out = os.path.join(HERE, "_build", "gpuAAK.c")
    os.makedirs(os.path.dirname(c_out), exist_ok=True)
    opts = Main.CompilationOptions(Main.default_options, output_file=c_out, language_level=3)
    res = Main.compile(pyx, opts)
    if res.num_errors:
        raise RuntimeError("cython failed on %s" % pyx)
    so = os.path.join(HERE, "gpuAAK" + sysconfig.get_config_var("EXT_SUFFIX"))
    cmd = ["gcc", "-shared", "-fPIC", "-O2", "-Wno-unused-function", "-I" + sysconfig.get_paths()["include"],
           "-I" + os.path.join(ROOT, "include"), c_out, "-L" + PKG, "-laak_hip", "-Wl,-rpath,$ORIGIN/..",
           "-o", so]
    subprocess.run(cmd, check=True)
    return so


if __name__ == "__main__":
    print(build())
    sys.exit(0)
