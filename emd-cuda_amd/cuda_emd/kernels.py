"""kernels.py -- EMDKernelConfig and EMDBatchLauncher over the gfx950 C ABI.

Same names, fields and call signatures as the reference's kernels.py
(SeismicFlow/emd-cuda kernels.py:1170-1407); the device code is the AOT-compiled HIP
library libemd_gfx950.so instead of an NVRTC-compiled CUDA string.  Arrays are torch
tensors on the current ROCm device (numpy inputs are copied there).

float64 only, exactly as the reference: EMDKernelConfig.dtype is a read-only property.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

# LDS per workgroup on gfx950 (MI355X); used when no device is queried.
GFX950_LDS_BYTES = 163840


def _torch():
    import torch
    return torch


@dataclass
class EMDKernelConfig:
    """Configuration of the batched sift (reference EMDKernelConfig, kernels.py:1170-1176).

    ``max_extrema`` is the per-side extrema cap: counts are clamped to it and a trace whose
    count reaches it is flagged in ``overflow``.  ``tile_slots`` (addition) is the LDS knot
    tile of the one-workgroup-per-trace launch; 0 sizes it from ``max_extrema`` like the
    reference's shared-memory budget.  A sift whose knots outgrow the LDS tile keeps them
    in the workgroup's global tile instead, so the tile never changes results.

    ``time_axis`` (addition) chooses what a caller's T means.  ``"index"`` (default):
    PyEMD's EMD() in its 'simple' extrema mode, where the sample index is the time axis
    whatever T is (a non-uniform T only warns).  ``"given"``: the reference kernel's use of
    T (kernels.py:471-472, 528-552, 703): knots sit at T[idx], mirrored knots at
    2 T[sym] - T[idx], and the envelopes are evaluated at T[i].  T is one axis [N] shared by
    the batch or one per signal [B,N] (the per-signal extension kernels.py:826-828 names);
    ``GPUBatchEMD`` normalises it as the reference does (emd.py:83-88).
    """

    max_extrema: int = 1024
    nbsym: int = 2
    max_iteration: int = 1000
    max_imfs: int = 16
    threads_per_block: int = 128
    tile_slots: int = 0
    time_axis: str = "index"

    def __post_init__(self):
        if self.time_axis not in ("index", "given"):
            raise ValueError(f"time_axis must be 'index' or 'given', got {self.time_axis!r}")

    @property
    def dtype(self) -> np.dtype:
        return np.dtype(np.float64)

    @property
    def real_t_name(self) -> str:
        return "double"

    @property
    def itemsize(self) -> int:
        return self.dtype.itemsize

    def static_shared_mem_bytes(self) -> int:
        """Static LDS of a sift workgroup (edge exchange + reduction scratch)."""
        return int(_lib.load().emd_static_lds_bytes(self.threads_per_block))

    def shared_mem_bytes(self) -> int:
        """Dynamic LDS of the knot tile: slots for both envelopes x bytes per slot
        (position int32; value, second derivative, 1/h fp64).  Reference counterpart:
        kernels.py:1273-1275."""
        slots = self.tile_slots or (2 * self.max_extrema + 4 * self.nbsym)
        return slots * int(_lib.load().emd_lds_slot_bytes())

    @classmethod
    def for_device(cls, device=None, target_blocks_per_sm: int = 4,
                   min_max_extrema: int = 1024, lds_bytes: int | None = None) -> "EMDKernelConfig":
        """Size ``max_extrema`` for occupancy like the reference (kernels.py:1205-1271):
        budget LDS for ``target_blocks_per_sm`` workgroups per CU, never below
        ``min_max_extrema``, and never beyond 95% of the per-workgroup LDS."""
        cfg = cls()
        if lds_bytes is None:
            lds = ctypes.c_int(0)
            if device is not None:
                torch = _torch()
                with torch.cuda.device(device):
                    _lib.check(_lib.load().emd_device_info(ctypes.byref(lds), None, None),
                               "EMDKernelConfig.for_device")
            else:
                _lib.check(_lib.load().emd_device_info(ctypes.byref(lds), None, None),
                           "EMDKernelConfig.for_device")
            lds_bytes = lds.value
        static_bytes = cfg.static_shared_mem_bytes()
        per_extremum = 2 * int(_lib.load().emd_lds_slot_bytes())  # one slot per envelope
        budget = max(0, lds_bytes // target_blocks_per_sm - static_bytes)
        occ = max(0, budget // per_extremum - 2 * cfg.nbsym)
        desired = max(min_max_extrema, occ)
        hard = max(64, (int(lds_bytes * 0.95) - static_bytes) // per_extremum - 2 * cfg.nbsym)
        cfg.max_extrema = min(desired, hard)
        return cfg

    def build(self) -> "EMDBatchLauncher":
        return EMDBatchLauncher(None, self)


def _as_device_f64(x, what: str):
    """torch float64 CUDA tensor from a torch tensor or numpy array; dtype is checked,
    never cast (reference: kernels.py:1346-1352)."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        if x.dtype != torch.float64:
            raise ValueError(
                f"This launcher was compiled for REAL_T=double (dtype float64), but got "
                f"{what}.dtype={x.dtype}. Passing mismatched dtypes would silently "
                f"reinterpret bytes rather than convert them -- cast explicitly first.")
        if not x.is_cuda:
            x = x.cuda()
        return x.contiguous()
    a = np.asarray(x)
    if a.dtype != np.float64:
        raise ValueError(
            f"This launcher was compiled for REAL_T=double (dtype float64), but got "
            f"{what}.dtype={a.dtype}. Passing mismatched dtypes would silently reinterpret "
            f"bytes rather than convert them -- cast explicitly first.")
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def warn_nonuniform_T(T) -> bool:
    """Warn (UserWarning) when T is not uniformly spaced; returns True if it warned.

    PyEMD's EMD() in its default 'simple' extrema mode replaces T by the sample index, and
    so does this build: a non-uniform T gives the decomposition of T = arange(N), not a
    time-warped one (the reference instead normalises T and uses it in its boundary mirror
    and spline evaluation, kernels.py:471-472, 528, 703; emd.py:83-88)."""
    import warnings
    T = np.asarray(T, dtype=np.float64)
    if T.size < 3:
        return False
    d = np.diff(T)
    if np.all(d > 0) and np.max(np.abs(d / d[0] - 1.0)) <= 1e-12:
        return False
    warnings.warn(
        "T is not uniformly spaced: like PyEMD's EMD() in its default 'simple' extrema "
        "mode, this build uses the sample index as the time axis, so the result is the "
        "decomposition for T = arange(N).", UserWarning, stacklevel=3)
    return True


def _ptr(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() > 0 else 0)


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


class EMDBatchLauncher:
    """Owns the library handle (reference EMDBatchLauncher, kernels.py:1289-1407).

    ``source`` is accepted for signature compatibility and ignored: the kernels are
    compiled ahead of time for gfx950.
    """

    def __init__(self, source=None, config: EMDKernelConfig | None = None):
        if config is None and isinstance(source, EMDKernelConfig):
            source, config = None, source
        self.config = config or EMDKernelConfig()
        lib = _lib.load()
        c = self.config
        cc = _lib.EmdConfigC(c.max_extrema, c.nbsym, c.max_iteration, c.max_imfs,
                             c.threads_per_block, c.tile_slots)
        h = ctypes.c_void_p()
        _lib.check(lib.emd_create(ctypes.byref(cc), ctypes.byref(h)), "EMDBatchLauncher")
        self._h = h
        self.last_stats = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().emd_destroy(h)
            except Exception:
                pass
            self._h = None

    # -- helpers ----------------------------------------------------------------------
    def _check_T(self, T, B, N):
        """Validate T.  Returns None (the sample index is the time axis) or, with
        time_axis='given', (device tensor of the times, per_signal flag)."""
        if T is None:
            return None
        torch = _torch()
        given = self.config.time_axis == "given"
        Tt = T if isinstance(T, torch.Tensor) else None
        Tn = np.asarray(T) if Tt is None else None
        dt = Tt.dtype if Tt is not None else Tn.dtype
        if dt not in (torch.float64, np.float64):
            raise ValueError(
                f"This launcher was compiled for REAL_T=double (dtype float64), but got "
                f"T.dtype={dt}.")
        shape = tuple(Tt.shape) if Tt is not None else Tn.shape
        if not given:
            if shape != (N,):
                raise ValueError(f"T must have shape ({N},), got {shape}")
            warn_nonuniform_T(Tt.detach().cpu().numpy() if Tt is not None else Tn)
            return None
        if shape not in ((N,), (B, N)):
            raise ValueError(f"T must have shape ({N},) or ({B}, {N}), got {shape}")
        dev = Tt.detach().to("cuda").contiguous() if Tt is not None else \
            torch.from_numpy(np.ascontiguousarray(Tn)).cuda()
        if N > 1 and not bool(torch.all(dev[..., 1:] > dev[..., :-1])):
            raise ValueError("T must increase strictly along each signal")
        return dev, len(shape) == 2

    def run(self, signals, T=None) -> tuple:
        """Decompose ``signals`` [B,N] float64.  Returns (imfs [B,max_imfs,N], n_imfs [B],
        overflow [B], debug_n, debug_extno, debug_nzm [B,max_imfs]) as torch tensors; the
        per-trace statistics [B,4] (sifts, peak extrema, knot sum, flags) are kept in
        ``last_stats``.  With the default time_axis='index' T is validated but, as in PyEMD's
        'simple' extrema mode, the sample index is the time axis; with time_axis='given' a
        T ([N] or [B,N], normalised as GPUBatchEMD does) places the knots and the envelope
        evaluation points, as in the reference kernel."""
        torch = _torch()
        s = _as_device_f64(signals, "signals")
        if s.dim() != 2:
            raise ValueError(f"signals must be [batch, N], got shape {tuple(s.shape)}")
        B, N = s.shape
        times = self._check_T(T, B, N)
        c = self.config
        dev = s.device
        imfs = torch.empty((B, c.max_imfs, N), dtype=torch.float64, device=dev)
        # every element below is written by the kernel (unreached debug slots as -1)
        i32 = dict(dtype=torch.int32, device=dev)
        n_imfs, overflow = torch.empty((B,), **i32), torch.empty((B,), **i32)
        dn, de, dz = (torch.empty((B, c.max_imfs), **i32) for _ in range(3))
        stats = torch.empty((B, 4), **i32)
        if times is None:
            _lib.check(_lib.load().emd_run_padded(
                self._h, _ptr(s), B, N, _ptr(imfs), _ptr(n_imfs), _ptr(overflow), _ptr(dn),
                _ptr(de), _ptr(dz), _ptr(stats), _stream()), "EMDBatchLauncher.run")
        else:
            _lib.check(_lib.load().emd_run_padded_times(
                self._h, _ptr(s), _ptr(times[0]), int(times[1]), B, N, _ptr(imfs),
                _ptr(n_imfs), _ptr(overflow), _ptr(dn), _ptr(de), _ptr(dz), _ptr(stats),
                _stream()), "EMDBatchLauncher.run")
        self.last_stats = stats
        return imfs, n_imfs, overflow, dn, de, dz

    def run_ragged(self, values, offsets, order=None, threads=None, tile_slots=0,
                   wgs_per_cu=0, persistent=True, out_offsets=None, times=None) -> tuple:
        """Decompose CSR traces (values, int64 offsets [B+1]).  ``order`` (int32 queue
        order) never changes outputs.  Returns (imfs flat, out_offsets [B] int64, n_imfs,
        overflow, debug_n, debug_extno, debug_nzm, stats); row k < n_imfs[s] of trace s is
        imfs[out_offsets[s] + k*len(s) : ... + len(s)] (rows past n_imfs are scratch).
        ``times`` (float64, laid out like values: each trace's normalised, strictly
        increasing time axis) runs the time-axis kernels whatever ``time_axis`` says."""
        torch = _torch()
        v = _as_device_f64(values, "values")
        tv = None
        if times is not None:
            tv = _as_device_f64(times, "times")
            if tv.numel() != v.numel():
                raise ValueError(f"times must have {v.numel()} entries (like values), got "
                                 f"{tv.numel()}")
        off = offsets if isinstance(offsets, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(offsets, dtype=np.int64)))
        if off.dtype != torch.int64:
            raise ValueError(f"offsets must be int64, got {off.dtype}")
        off_host = off.cpu().numpy()
        off = off.cuda().contiguous()
        B = int(off.numel()) - 1
        if B < 0:
            raise ValueError("offsets needs at least one entry")
        lengths = np.diff(off_host)
        if B and (off_host[0] < 0 or lengths.min() < 0 or off_host[-1] > v.numel()):
            raise ValueError("offsets must start at >= 0, be non-decreasing and stay "
                             "within values")
        max_n = int(lengths.max()) if B else 0
        c = self.config
        dev = v.device
        if out_offsets is None:
            out_offsets = off[:-1] * c.max_imfs
            total = int(off_host[-1]) * c.max_imfs
        else:
            out_offsets = out_offsets.cuda().contiguous()
            total = int((out_offsets.cpu().numpy() + c.max_imfs * lengths).max()) if B else 0
        imfs = torch.empty((max(total, 1),), dtype=torch.float64, device=dev)
        # every element below is written by the kernel (unreached debug slots as -1)
        i32 = dict(dtype=torch.int32, device=dev)
        n_imfs, overflow = torch.empty((B,), **i32), torch.empty((B,), **i32)
        dn, de, dz = (torch.empty((B, c.max_imfs), **i32) for _ in range(3))
        stats = torch.empty((B, 4), **i32)
        if order is not None:
            if not (isinstance(order, torch.Tensor) and order.is_cuda):
                # a host order is checked here: a permutation of the B traces (an entry
                # out of range would be skipped by the kernel, a repeated one would leave
                # another trace's outputs unwritten).  A device order is the caller's
                # promise; the kernel still skips entries outside [0, B).
                oh = np.asarray(order.numpy() if isinstance(order, torch.Tensor) else order)
                if oh.shape != (B,):
                    raise ValueError(f"order must have {B} entries, got shape {oh.shape}")
                if B and (oh.min() < 0 or oh.max() >= B
                          or not np.all(np.bincount(oh.astype(np.int64), minlength=B) == 1)):
                    raise ValueError("order must be a permutation of range(B)")
            order = (order if isinstance(order, torch.Tensor) else torch.from_numpy(
                np.ascontiguousarray(np.asarray(order, dtype=np.int32)))).to(dev).int().contiguous()
            if order.numel() != B:
                raise ValueError(f"order must have {B} entries, got {order.numel()}")
        L = _lib.EmdLaunchC(threads or c.threads_per_block, tile_slots, wgs_per_cu,
                            1 if persistent else 0, B)
        if tv is None:
            _lib.check(_lib.load().emd_run_ragged(
                self._h, _ptr(v), _ptr(off), B, max_n, _ptr(order), ctypes.byref(L),
                _ptr(imfs), _ptr(out_offsets), _ptr(n_imfs), _ptr(overflow), _ptr(dn),
                _ptr(de), _ptr(dz), _ptr(stats), _stream()), "EMDBatchLauncher.run_ragged")
        else:
            _lib.check(_lib.load().emd_run_ragged_times(
                self._h, _ptr(v), _ptr(tv), _ptr(off), B, max_n, _ptr(order), ctypes.byref(L),
                _ptr(imfs), _ptr(out_offsets), _ptr(n_imfs), _ptr(overflow), _ptr(dn),
                _ptr(de), _ptr(dz), _ptr(stats), _stream()), "EMDBatchLauncher.run_ragged")
        self.last_stats = stats
        return imfs, out_offsets, n_imfs, overflow, dn, de, dz, stats

    def run_envelope_only(self, signals, T=None) -> tuple:
        """Extrema + mirror + both envelope splines of the raw signals, no sifting
        (reference kernels.py:1374-1407).  Returns (max_env, min_env, n_extrema [B,2],
        degenerate [B], max_knot_pos, max_knot_val, min_knot_pos, min_knot_val [B,MBUF],
        n_mirrored [B,2]).  T as in ``run`` (knot positions in time units with
        time_axis='given')."""
        torch = _torch()
        s = _as_device_f64(signals, "signals")
        if s.dim() != 2:
            raise ValueError(f"signals must be [batch, N], got shape {tuple(s.shape)}")
        B, N = s.shape
        times = self._check_T(T, B, N)
        c = self.config
        mbuf = c.max_extrema + 2 * c.nbsym
        dev = s.device
        z = lambda *shape, dt=torch.float64: torch.zeros(shape, dtype=dt, device=dev)
        max_env, min_env = z(B, N), z(B, N)
        n_extrema, degenerate = z(B, 2, dt=torch.int32), z(B, dt=torch.int32)
        mkp, mkv, nkp, nkv = z(B, mbuf), z(B, mbuf), z(B, mbuf), z(B, mbuf)
        n_mirrored = z(B, 2, dt=torch.int32)
        if times is None:
            _lib.check(_lib.load().emd_envelope_only(
                self._h, _ptr(s), B, N, _ptr(max_env), _ptr(min_env), _ptr(n_extrema),
                _ptr(degenerate), _ptr(mkp), _ptr(mkv), _ptr(nkp), _ptr(nkv),
                _ptr(n_mirrored), _stream()), "EMDBatchLauncher.run_envelope_only")
        else:
            _lib.check(_lib.load().emd_envelope_only_times(
                self._h, _ptr(s), _ptr(times[0]), int(times[1]), B, N, _ptr(max_env),
                _ptr(min_env), _ptr(n_extrema), _ptr(degenerate), _ptr(mkp), _ptr(mkv),
                _ptr(nkp), _ptr(nkv), _ptr(n_mirrored), _stream()),
                "EMDBatchLauncher.run_envelope_only")
        return max_env, min_env, n_extrema, degenerate, mkp, mkv, nkp, nkv, n_mirrored

    def compact(self, imfs, n_imfs, offsets=None, N: int | None = None, out_offsets=None):
        """Keep each trace's n_imfs rows, back to back: returns (row_offsets [B+1] int64,
        compact fp64 [row_offsets[B]]); trace s's rows start at row_offsets[s].  ``imfs``
        is run()'s padded output (give N) or run_ragged()'s flat output (give offsets)."""
        torch = _torch()
        c = self.config
        B = int(n_imfs.numel())
        dev = n_imfs.device
        off = None if offsets is None else (
            offsets if isinstance(offsets, torch.Tensor) else torch.from_numpy(
                np.asarray(offsets, dtype=np.int64))).to(dev).contiguous()
        if off is None and N is None:
            N = int(imfs.shape[-1])
        row = torch.empty((B + 1,), dtype=torch.int64, device=dev)
        lib = _lib.load()
        _lib.check(lib.emd_compact_offsets(_ptr(n_imfs), _ptr(off), int(N or 0), B, _ptr(row),
                                           _stream()), "compact")
        total = int(row[B].item()) if B else 0
        out = torch.empty((max(total, 1),), dtype=torch.float64, device=dev)
        _lib.check(lib.emd_compact(_ptr(imfs), _ptr(out_offsets), c.max_imfs, _ptr(n_imfs),
                                   _ptr(off), int(N or 0), B, _ptr(row), _ptr(out), _stream()),
                   "compact")
        return row, out[:total]

    def set_timing(self, enable: bool = True) -> None:
        """Record HIP events on the launch stream around the sift kernel launches of every
        following run / run_ragged call."""
        _lib.check(_lib.load().emd_set_timing(self._h, 1 if enable else 0), "set_timing")

    def kernel_time_ms(self) -> float:
        """Milliseconds between the events of the last timed run (waits for it)."""
        ms = ctypes.c_float(0.0)
        _lib.check(_lib.load().emd_kernel_time_ms(self._h, ctypes.byref(ms)), "kernel_time_ms")
        return float(ms.value)

    def max_resident(self, threads: int, tile_slots: int, max_n: int, persistent=True) -> int:
        L = _lib.EmdLaunchC(threads, tile_slots, 0, 1 if persistent else 0)
        out = ctypes.c_int(0)
        _lib.check(_lib.load().emd_max_resident(self._h, ctypes.byref(L), max_n,
                                                 ctypes.byref(out)), "max_resident")
        return out.value


class ChunkGraph:
    """One decomposition over fixed device buffers, captured once into a HIP graph.

    Buffers (allocated here) hold up to ``capacity`` traces and ``values_capacity``
    samples in CSR form (a padded chunk is CSR with uniform offsets); every trace must be
    at most ``max_n`` long.  Refill ``values`` / ``offsets`` (and ``order``) between
    replays -- ``load()`` does it on the current stream -- and ``launch()`` replays the
    sift, then the compaction (``compact=True``) into ``row_offsets`` / ``compact_out``.
    Unused trace slots are zero-length traces.  Outputs follow run_ragged's layout.
    """

    def __init__(self, launcher: "EMDBatchLauncher", capacity: int, values_capacity: int,
                 max_n: int, threads: int | None = None, tile_slots: int = 0,
                 wgs_per_cu: int = 0, persistent: bool = True, compact: bool = True,
                 use_order: bool = False):
        torch = _torch()
        c = launcher.config
        self.launcher = launcher
        self.capacity, self.values_capacity, self.max_n = capacity, values_capacity, max_n
        dev = torch.device("cuda", torch.cuda.current_device())
        i32 = dict(dtype=torch.int32, device=dev)
        self.values = torch.zeros((max(values_capacity, 1),), dtype=torch.float64, device=dev)
        self.offsets = torch.zeros((capacity + 1,), dtype=torch.int64, device=dev)
        self.order = torch.arange(capacity, **i32) if use_order else None
        self.imfs = torch.empty((max(values_capacity * c.max_imfs, 1),), dtype=torch.float64,
                                device=dev)
        self.n_imfs, self.overflow = torch.empty((capacity,), **i32), torch.empty((capacity,), **i32)
        self.debug_n, self.debug_extno, self.debug_nzm = (
            torch.empty((capacity, c.max_imfs), **i32) for _ in range(3))
        self.stats = torch.empty((capacity, 4), **i32)
        self.row_offsets = torch.empty((capacity + 1,), dtype=torch.int64, device=dev) \
            if compact else None
        self.compact_out = torch.empty((max(values_capacity * c.max_imfs, 1),),
                                       dtype=torch.float64, device=dev) if compact else None
        L = _lib.EmdLaunchC(threads or c.threads_per_block, tile_slots, wgs_per_cu,
                            1 if persistent else 0)
        spec = _lib.EmdChunkC(_ptr(self.values).value, _ptr(self.offsets).value, 0, capacity,
                              max_n, _ptr(self.order).value, L, _ptr(self.imfs).value, None,
                              _ptr(self.n_imfs).value, _ptr(self.overflow).value,
                              _ptr(self.debug_n).value, _ptr(self.debug_extno).value,
                              _ptr(self.debug_nzm).value, _ptr(self.stats).value,
                              _ptr(self.row_offsets).value, _ptr(self.compact_out).value)
        torch.cuda.synchronize()
        g = ctypes.c_void_p()
        _lib.check(_lib.load().emd_chunk_graph_create(launcher._h, ctypes.byref(spec),
                                                       ctypes.byref(g)), "ChunkGraph")
        self._g = g
        self.count = 0

    def load(self, values, offsets, order=None, non_blocking: bool = True) -> None:
        """Copy one chunk (values, int64 offsets [B+1] starting at 0, B <= capacity) into
        the buffers on the current stream; host tensors should be pinned."""
        torch = _torch()
        off = offsets if isinstance(offsets, torch.Tensor) else torch.from_numpy(
            np.asarray(offsets, dtype=np.int64))
        B = off.numel() - 1
        if B > self.capacity:
            raise ValueError(f"chunk of {B} traces exceeds the capacity {self.capacity}")
        n_vals = int(off[-1]) if off.device.type == "cpu" else int(off[-1].item())
        if n_vals > self.values_capacity:
            raise ValueError(f"chunk of {n_vals} samples exceeds {self.values_capacity}")
        if off.device.type == "cpu":
            lens = np.diff(off.numpy())
            if B and int(lens.max()) > self.max_n:
                raise ValueError(f"a trace of {int(lens.max())} samples exceeds max_n "
                                 f"{self.max_n}")
        v = values if isinstance(values, torch.Tensor) else torch.from_numpy(np.asarray(values))
        self.values[:n_vals].copy_(v.reshape(-1)[:n_vals], non_blocking=non_blocking)
        self.offsets[:B + 1].copy_(off, non_blocking=non_blocking)
        self.offsets[B + 1:].fill_(n_vals)
        if self.order is not None:
            if order is None:
                self.order.copy_(torch.arange(self.capacity, dtype=torch.int32))
            else:
                o = order if isinstance(order, torch.Tensor) else torch.from_numpy(
                    np.asarray(order, dtype=np.int32))
                self.order[:B].copy_(o, non_blocking=non_blocking)
                self.order[B:].copy_(torch.arange(B, self.capacity, dtype=torch.int32))
        self.count = B

    def launch(self) -> None:
        _lib.check(_lib.load().emd_chunk_graph_launch(self._g, _stream()), "ChunkGraph.launch")

    def __del__(self):
        g = getattr(self, "_g", None)
        if g is not None and g.value:
            try:
                _lib.load().emd_chunk_graph_destroy(g)
            except Exception:
                pass
            self._g = None
