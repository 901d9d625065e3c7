"""emd.py -- GPUBatchEMD, the user-facing wrapper (reference emd.py:53-114).

float64 only.  Usage (same as the reference, with torch tensors instead of CuPy arrays):

    from cuda_emd.emd import GPUBatchEMD
    gpu_emd = GPUBatchEMD()
    gpu_emd.warmup()
    imfs, n_imfs, overflow = gpu_emd(signals)      # signals: [batch, N] float64
    imfs_i = gpu_emd.unpack(imfs, n_imfs, i=0)     # numpy (n_imfs[i], N), like PyEMD EMD()(x)

Outputs: imfs [batch, MAX_IMFS, N] float64, n_imfs [batch] int32, overflow [batch] int32
(1 = MAX_EXTREMA or MAX_ITERATION was hit: the decomposition is unverified);
``last_debug`` = (debug_n, debug_extno, debug_nzm) [batch, MAX_IMFS] int32.

Warmup: the kernels are compiled ahead of time, so there is no JIT; warmup() still runs
one tiny call so that library loading, device initialisation and first-launch costs stay
out of timed regions.
"""
from __future__ import annotations

import warnings

import numpy as np

try:
    from .kernels import EMDKernelConfig, EMDBatchLauncher
except ImportError:  # flat layout: emd.py and kernels.py side by side
    from kernels import EMDKernelConfig, EMDBatchLauncher  # type: ignore


class GPUBatchEMD:
    def __init__(self, config: "EMDKernelConfig | None" = None):
        self._config = config or EMDKernelConfig.for_device()
        self._launcher = None
        self.last_debug = None
        self.last_stats = None

    def warmup(self, N: int = 64) -> None:
        """One tiny real call (all-zero trace) before timing anything."""
        import torch
        self(torch.zeros((1, N), dtype=torch.float64, device="cuda"))

    def _get_launcher(self) -> EMDBatchLauncher:
        if self._launcher is None:
            self._launcher = self._config.build()
        return self._launcher

    def __call__(self, signals, T=None):
        import torch
        if not isinstance(signals, torch.Tensor):
            signals = np.asarray(signals)
        if signals.ndim == 1:
            signals = signals[None, :]
        dt = signals.dtype
        if dt not in (torch.float64, np.float64):
            raise ValueError(
                f"GPUBatchEMD (this build) is float64 only, got {dt}. Cast explicitly first, "
                f"e.g. signals.astype(np.float64) / signals.double() -- this class won't "
                f"silently do that for you.")
        launcher = self._get_launcher()
        batch, N = signals.shape
        if T is not None:
            # the reference's normalisation (emd.py:83-88), per signal for a [B,N] axis
            Tn = T.detach().cpu().numpy() if hasattr(T, "detach") else np.asarray(T)
            if Tn.dtype != np.float64:
                raise ValueError(f"T must be float64, got {Tn.dtype}")
            if Tn.shape[-1:] != (N,) or Tn.ndim > 2:
                raise ValueError(f"T must have shape ({N},) or ({batch}, {N}), got {Tn.shape}")
            if N > 1:
                d = np.diff(Tn, axis=-1)
                T = (Tn - Tn[..., :1]) / np.min(d, axis=-1, keepdims=True)
            else:
                T = Tn - Tn[..., :1]
        imfs, n_imfs, overflow, debug_n, debug_extno, debug_nzm = launcher.run(signals, T)
        self.last_debug = (debug_n, debug_extno, debug_nzm)
        self.last_stats = launcher.last_stats
        if bool(torch.any(overflow)):
            n_hit = int(torch.sum(overflow))
            warnings.warn(
                f"{n_hit}/{batch} signal(s) hit MAX_EXTREMA or MAX_ITERATION during "
                "sifting -- their output is not a verified decomposition. Check the "
                "`overflow` array to find which, and raise EMDKernelConfig.max_extrema "
                "/ max_iteration for those.",
                stacklevel=2,
            )
        return imfs, n_imfs, overflow

    def unpack(self, imfs, n_imfs, i: int) -> np.ndarray:
        """Signal i's IMFs (residue as the last row) trimmed to n_imfs[i], as numpy
        (n_imfs[i], N) -- directly comparable to PyEMD's EMD()(signal)."""
        k = int(n_imfs[i])
        return imfs[i, :k, :].detach().cpu().numpy()
