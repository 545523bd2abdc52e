import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def _build_missing():
    """A tree with nothing built (a fresh checkout) is built once, as
    __graft_entry__.build() does; a built tree is used as it is."""
    import subprocess
    libs = {os.path.join(ROOT, "emri_kludge_suite_amd", "csrc"): os.path.join(ROOT, "emri_kludge_suite_amd", "libaak_hip.so"),
            os.path.join(ROOT, "oracle"): os.path.join(ROOT, "oracle", "liboracle.so")}
    jobs = str(min(8, os.cpu_count() or 1))
    for src, lib in libs.items():
        if not os.path.exists(lib):
            subprocess.run(["make", "-C", src, "-j", jobs], check=True)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")
    _build_missing()
