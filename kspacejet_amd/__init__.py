"""KSpaceJet reconstruction hot path for AMD Instinct MI355X (gfx950).

Native code lives in libksj_gpu.so (HIP kernels + the C ABI of include/ksj_gpu.h)
and the KSpaceJet Provider modules libksj_provider_<module>.so (noncartesian_recon,
cartesian_recon, coil_combine, kspace_conditioning, calibration, gpu_recon).  Python
here is plumbing: ctypes bindings, torch-allocated device buffers and synthetic data.
"""
from . import _lib  # noqa: F401

__all__ = ["_lib"]
