"""MI355X-native AAK waveform + LISA likelihood (drop-in for the EMRI Kludge
Suite's GPU path: pygpuAAK's GPUAAK class and the gpuAAK.h entry points).

The compute runs in the in-tree HIP library libaak_hip.so (hand-written
gfx950 kernels + rocFFT) behind the C ABI declared in include/gpuAAK.h.

  GPUAAK       the reference's GPUAAK class (plus LikelihoodBatch)
  OmegaPhi, PNevolution   the free functions of the reference's gpuAAK.h
  pyGPUAAK     the reference's likelihood caller, batched
  AAKwrapper   wave() and phase() of the reference's CPU wrapper, on the GPU
  Converter    sampler-to-physical parameter conversion
"""
from . import AAKwrapper  # noqa: F401
from .convert import Converter  # noqa: F401
from .gpuAAK import GPUAAK, OmegaPhi, PNevolution  # noqa: F401
from .pygpuAAK import pyGPUAAK  # noqa: F401

__all__ = ["GPUAAK", "OmegaPhi", "PNevolution", "pyGPUAAK", "AAKwrapper", "Converter"]
