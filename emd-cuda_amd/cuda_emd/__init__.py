"""cuda_emd -- batched PyEMD-default Empirical Mode Decomposition on MI355X (gfx950).

Drop-in for SeismicFlow/emd-cuda's ``cuda_emd`` package: ``GPUBatchEMD`` (emd.py),
``EMDKernelConfig`` and ``EMDBatchLauncher`` (kernels.py), backed by hand-written HIP
kernels in libemd_gfx950.so.
"""
from .kernels import EMDBatchLauncher, EMDKernelConfig
from .emd import GPUBatchEMD

__all__ = ["GPUBatchEMD", "EMDKernelConfig", "EMDBatchLauncher"]
