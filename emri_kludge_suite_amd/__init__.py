"""MI355X-native AAK waveform + LISA likelihood (drop-in for the EMRI Kludge
Suite's GPU path: pygpuAAK's GPUAAK class and the gpuAAK.h entry points).

The compute runs in the in-tree HIP library libaak_hip.so (hand-written
gfx950 kernels + rocFFT) behind the C ABI declared in include/gpuAAK.h.
"""
from .gpuAAK import GPUAAK, OmegaPhi  # noqa: F401

__all__ = ["GPUAAK", "OmegaPhi"]
