"""hilbert.py -- the Hilbert spectral step over IMF rows (``emd_hilbert`` in libemd_gfx950.so).

Instantaneous amplitude, phase and frequency of every IMF row, from the row's analytic
signal z = x + i H x with H x as ``scipy.signal.hilbert`` computes it:

- ``n_fft=0`` (default): ``hilbert(row)``, a DFT of the row's own length (rows of up to
  8192 samples, power-of-two rows up to 16384);
- ``n_fft=-1``: ``hilbert(row, N=L)[:len]`` with L the smallest power of two >= len;
- ``n_fft=L`` (a power of two >= the longest row, <= 16384): the same L for every row.

``amp = |z|``, ``phase = numpy.angle(z)``, and ``freq[n] = diff(unwrap(angle(z)))[n] / (2 pi
dt)`` for n < len-1 with the last sample repeating the one before (0 for a one-sample row),
so with ``dt`` the sample interval in seconds (0.004 for 4 ms traces) ``freq`` is in Hz.

Layouts (the outputs share the input's):

- padded ``[B, max_imfs, N]`` (``GPUBatchEMD`` / ``EMDBatchLauncher.run``): rows past
  ``n_imfs`` come back as zeros;
- flat rows of CSR traces, given ``offsets`` (int64 [B+1]): ``run_ragged``'s output with its
  ``out_offsets`` as ``row_base`` (or None for the default ``offsets[s] * max_imfs``), or
  ``compact``'s output with its ``row_offsets`` as ``row_base``; elements outside the rows
  are left as allocated (uninitialised).

No reference counterpart: the reference stops at the IMFs (kernels.py:1343-1372) of
traces it describes as seismic, sampled at 4 ms (emd.py:28-30).
"""
from __future__ import annotations

import numpy as np

from . import _lib
from .kernels import _ptr, _stream

OUTPUTS = ("amp", "phase", "freq")


def _torch():
    import torch
    return torch


class HilbertStep:
    """The step over fixed buffers: the layout is checked and the outputs are allocated
    once; every ``run()`` launches on the current stream against ``imfs``' current contents
    and returns {name: tensor} for the requested ``outputs``, each shaped like ``imfs``."""

    def __init__(self, imfs, n_imfs, offsets=None, row_base=None, max_imfs: int | None = None,
                 dt: float = 1.0, n_fft: int = 0, outputs=OUTPUTS):
        torch = _torch()
        want = tuple(outputs)
        bad = [o for o in want if o not in OUTPUTS]
        if bad or not want:
            raise ValueError(f"outputs must be a non-empty subset of {OUTPUTS}, got {want}")
        if not isinstance(imfs, torch.Tensor) or imfs.dtype != torch.float64:
            raise ValueError("imfs must be a float64 torch tensor")
        dev = imfs.device
        if dev.type != "cuda":
            raise ValueError("imfs must be on the GPU")
        if not imfs.is_contiguous():
            raise ValueError("imfs must be contiguous")
        self.imfs = imfs
        self.n_imfs = torch.as_tensor(n_imfs).to(device=dev, dtype=torch.int32).contiguous()
        B = int(self.n_imfs.numel())
        if offsets is None:
            if imfs.dim() != 3 or imfs.shape[0] != B:
                raise ValueError(f"padded imfs must be [B={B}, max_imfs, N], got "
                                 f"{tuple(imfs.shape)}")
            if row_base is not None:
                raise ValueError("row_base applies to flat rows with offsets")
            K, N = int(imfs.shape[1]), int(imfs.shape[2])
            if B and int(self.n_imfs.max()) > K:
                raise ValueError(f"n_imfs exceeds the {K} rows of imfs")
            self.offsets, self.row_base, self.max_n, self.max_imfs, self.N = None, None, N, K, N
            self.out = {o: torch.zeros_like(imfs) for o in want}
        else:
            off_h = np.asarray(offsets.cpu().numpy() if isinstance(offsets, torch.Tensor)
                               else offsets, dtype=np.int64)
            if off_h.shape != (B + 1,):
                raise ValueError(f"offsets must have {B + 1} entries, got {off_h.shape}")
            lengths = np.diff(off_h)
            if B and (off_h[0] < 0 or lengths.min() < 0):
                raise ValueError("offsets must start at >= 0 and be non-decreasing")
            self.max_n = int(lengths.max()) if B else 0
            self.N = 0
            nim_h = self.n_imfs.cpu().numpy().astype(np.int64)
            if row_base is None:
                mk = int(max_imfs) if max_imfs is not None else 0
                if mk < 1:
                    raise ValueError("the default flat layout needs max_imfs")
                if B and nim_h.max() > mk:
                    raise ValueError(f"n_imfs exceeds max_imfs {mk}")
                ends = off_h[:-1] * mk + nim_h * lengths
                self.row_base = None
            else:
                rb_h = np.asarray(row_base.cpu().numpy() if isinstance(row_base, torch.Tensor)
                                  else row_base, dtype=np.int64)[:B]
                if rb_h.shape != (B,):
                    raise ValueError(f"row_base needs {B} entries")
                if B and rb_h.min() < 0:
                    raise ValueError("row_base entries must be >= 0")
                ends = rb_h + nim_h * lengths
                mk = int(max_imfs or 1)
                self.row_base = torch.from_numpy(np.ascontiguousarray(rb_h)).to(dev)
            if B and int(ends.max()) > imfs.numel():
                raise ValueError("the rows described reach past the end of imfs")
            self.max_imfs = mk
            self.min_n = int(lengths.min()) if B else 0
            if B and self.min_n == self.max_n:
                # one length: the rows' bases make the lengths' offsets unnecessary, and the
                # library launches one workgroup per trace in a single launch
                if self.row_base is None:
                    self.row_base = torch.from_numpy(off_h[:-1] * mk).to(dev)
                self.offsets, self.N = None, self.max_n
            else:
                self.offsets = torch.from_numpy(off_h).to(dev)
            self.out = {o: torch.empty_like(imfs) for o in want}
        self.B, self.n_fft, self.dt = B, int(n_fft), float(dt)

    def run(self) -> dict:
        _lib.check(_lib.load().emd_hilbert(
            _ptr(self.imfs), _ptr(self.row_base), self.max_imfs, _ptr(self.n_imfs),
            _ptr(self.offsets), self.N, self.B, getattr(self, "min_n", self.max_n), self.max_n,
            self.n_fft, self.dt,
            _ptr(self.out.get("amp")), _ptr(self.out.get("phase")), _ptr(self.out.get("freq")),
            _stream()), "hilbert_spectrum")
        return self.out


def hilbert_spectrum(imfs, n_imfs, offsets=None, row_base=None, max_imfs: int | None = None,
                     dt: float = 1.0, n_fft: int = 0, outputs=OUTPUTS) -> dict:
    """Returns {name: tensor} for the requested ``outputs`` (a subset of amp, phase, freq),
    each shaped like ``imfs``, computed on ``imfs``' device on the current stream."""
    _torch()
    if hasattr(imfs, "contiguous"):
        imfs = imfs.contiguous()
    return HilbertStep(imfs, n_imfs, offsets, row_base, max_imfs, dt, n_fft, outputs).run()
