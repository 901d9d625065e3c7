"""traceio.py -- the data formats either side of the decomposition.

- SEG-Y in: ``scan_segy`` reads the binary file header and walks the 240-byte trace
  headers; ``read_segy`` copies the trace bytes to the device once and decodes every sample
  there (``emd_segy_decode``) to the float64 rows ``GPUBatchEMD`` / ``run_ragged`` take:
  a [B, N] batch when every trace has the file's sample count, else a CSR batch
  (values + int64 offsets).
  ``iter_segy`` does the same chunk by chunk for files larger than memory.
- CSR batches and compacted IMF rows on disk: ``save_csr`` / ``load_csr`` and
  ``save_compacted`` / ``load_compacted`` (``.npz``, no pickled objects).

No reference counterpart: the reference takes float64 device arrays (emd.py:60-69) of
seismic traces sampled at 4 ms (emd.py:28-30).  Header positions follow the SEG-Y rev 1
standard (1-based byte numbers): sample interval in microseconds at 3217-3218, samples per
trace at 3221-3222, sample format code at 3225-3226, extended textual header count at
3505-3506 (all big-endian), and a trace's own sample count at bytes 115-116 of its header.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .kernels import _ptr, _stream

TEXT_HEADER_BYTES = 3200
BINARY_HEADER_BYTES = 400
TRACE_HEADER_BYTES = 240
SAMPLE_BYTES = {1: 4, 2: 4, 3: 2, 5: 4, 6: 8, 8: 1}


def _torch():
    import torch
    return torch


def _be16(buf: np.ndarray, pos: int, signed: bool = False) -> int:
    v = (int(buf[pos]) << 8) | int(buf[pos + 1])
    return v - 0x10000 if signed and v >= 0x8000 else v


@dataclass
class SegyLayout:
    """Where the traces of one SEG-Y file lie.  ``trace_pos`` is relative to
    ``data_start`` (the first trace header); ``lengths`` are samples per trace."""
    format: int
    sample_bytes: int
    dt_us: int
    n_samples: int
    data_start: int
    trace_pos: np.ndarray
    lengths: np.ndarray
    fixed: bool

    @property
    def n_traces(self) -> int:
        return int(self.lengths.size)

    @property
    def stride(self) -> int:
        return TRACE_HEADER_BYTES + self.n_samples * self.sample_bytes

    @property
    def offsets(self) -> np.ndarray:
        return np.concatenate([[0], np.cumsum(self.lengths)]).astype(np.int64)


def scan_segy(buf: np.ndarray) -> SegyLayout:
    """Parse the file headers and locate every trace (``buf``: the whole file as uint8)."""
    buf = np.asarray(buf, dtype=np.uint8)
    size = buf.size
    if size < TEXT_HEADER_BYTES + BINARY_HEADER_BYTES:
        raise ValueError(f"{size} bytes is shorter than the 3600-byte SEG-Y file header")
    fmt = _be16(buf, 3224)
    if fmt not in SAMPLE_BYTES:
        raise ValueError(f"unsupported SEG-Y sample format code {fmt} "
                         "(supported: 1, 2, 3, 5, 6, 8)")
    s = SAMPLE_BYTES[fmt]
    dt_us = _be16(buf, 3216)
    ns = _be16(buf, 3220)
    n_ext = _be16(buf, 3504, signed=True)
    if n_ext < 0:
        raise ValueError("a variable number of extended textual headers (-1) is not supported")
    start = TEXT_HEADER_BYTES + BINARY_HEADER_BYTES + TEXT_HEADER_BYTES * n_ext
    body = size - start
    if body < 0:
        raise ValueError("the file ends inside its extended textual headers")
    stride = TRACE_HEADER_BYTES + ns * s
    if body % stride == 0:
        B = body // stride
        pos = np.arange(B, dtype=np.int64) * stride
        if B:
            hdr = buf[start:].reshape(B, stride)[:, 114:116].astype(np.int64)
            own = (hdr[:, 0] << 8) | hdr[:, 1]
            if np.all((own == ns) | (own == 0)):
                return SegyLayout(fmt, s, dt_us, ns, start, pos,
                                  np.full(B, ns, dtype=np.int64), True)
        else:
            return SegyLayout(fmt, s, dt_us, ns, start, pos, np.zeros(0, np.int64), True)
    # Traces of their own lengths: walk the headers.
    pos, lens = [], []
    p = 0
    while p < body:
        if p + TRACE_HEADER_BYTES > body:
            raise ValueError(f"trace header at byte {start + p} runs past the end of the file")
        n = _be16(buf, start + p + 114) or ns
        end = p + TRACE_HEADER_BYTES + n * s
        if end > body:
            raise ValueError(f"trace at byte {start + p} ({n} samples) runs past the end "
                             "of the file")
        pos.append(p)
        lens.append(n)
        p = end
    return SegyLayout(fmt, s, dt_us, ns, start, np.asarray(pos, np.int64),
                      np.asarray(lens, np.int64), False)


@dataclass
class SegyTraces:
    """Decoded traces on the device: ``signals`` [B, N] float64 when every trace has the
    file's sample count, else ``values`` + ``offsets`` (CSR).  ``dt`` is in seconds."""
    layout: SegyLayout
    dt: float
    signals: object = None
    values: object = None
    offsets: object = None


def decode_segy(dev_bytes, layout: SegyLayout, out=None):
    """Decode the traces held in ``dev_bytes`` (a uint8 device tensor starting at the first
    trace header) on the current stream.  Returns the [B, N] rows for a fixed-length
    layout, else (values, offsets)."""
    torch = _torch()
    lib = _lib.load()
    if dev_bytes.dtype != torch.uint8 or not dev_bytes.is_cuda:
        raise ValueError("dev_bytes must be a uint8 device tensor")
    need = int(layout.trace_pos[-1] + TRACE_HEADER_BYTES
               + layout.lengths[-1] * layout.sample_bytes) if layout.n_traces else 0
    if dev_bytes.numel() < need:
        raise ValueError(f"dev_bytes holds {dev_bytes.numel()} bytes; the layout needs {need}")
    B = layout.n_traces
    dev = dev_bytes.device
    if layout.fixed:
        N = layout.n_samples
        if out is None:
            out = torch.empty((B, N), dtype=torch.float64, device=dev)
        elif tuple(out.shape) != (B, N) or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f"out must be a contiguous float64 [{B}, {N}] tensor")
        _lib.check(lib.emd_segy_decode(_ptr(dev_bytes), None, ctypes.c_int64(layout.stride),
                                       TRACE_HEADER_BYTES, None, N, B, layout.format,
                                       _ptr(out), _stream()), "emd_segy_decode")
        return out
    if out is not None:
        raise ValueError("out is for fixed-length layouts; a ragged layout allocates its "
                         "CSR values")
    offsets_h = layout.offsets
    if int(layout.lengths.max(initial=0)) >= 2 ** 31:
        raise ValueError("a trace longer than 2^31 - 1 samples")
    offsets = torch.from_numpy(offsets_h).to(dev)
    pos = torch.from_numpy(layout.trace_pos).to(dev)
    values = torch.empty(int(offsets_h[-1]), dtype=torch.float64, device=dev)
    _lib.check(lib.emd_segy_decode(_ptr(dev_bytes), _ptr(pos), ctypes.c_int64(0),
                                   TRACE_HEADER_BYTES, _ptr(offsets), 0, B, layout.format,
                                   _ptr(values), _stream()), "emd_segy_decode")
    return values, offsets


def read_segy(path: str, device=None) -> SegyTraces:
    """Read a SEG-Y file: headers on the host, one pinned copy of the trace bytes to the
    device, every sample decoded there."""
    torch = _torch()
    layout = scan_segy(np.memmap(path, dtype=np.uint8, mode="r"))
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    # The trace bytes go straight from the file into pinned memory (one host copy).
    with open(path, "rb") as f:
        f.seek(layout.data_start)
        host = torch.empty(os.fstat(f.fileno()).st_size - layout.data_start,
                           dtype=torch.uint8, pin_memory=True)
        view = memoryview(host.numpy())
        got = 0
        while got < len(view):
            n = f.readinto(view[got:])
            if not n:
                raise ValueError(f"{path} shrank while being read")
            got += n
    dev_bytes = host.to(dev, non_blocking=True)
    res = decode_segy(dev_bytes, layout)
    torch.cuda.current_stream(dev).synchronize()
    out = SegyTraces(layout=layout, dt=layout.dt_us * 1e-6)
    if layout.fixed:
        out.signals = res
    else:
        out.values, out.offsets = res
    return out


def iter_segy(path: str, chunk: int = 16384, device=None):
    """Read a SEG-Y file chunk by chunk (files larger than host or device memory): yields
    (first trace index, SegyTraces) for each run of up to ``chunk`` traces, in file order.
    The file is memory-mapped; each chunk's bytes are staged through one of two pinned
    buffers and copied and decoded on the current stream, so filling the next chunk on the
    host overlaps the device's copy and decode of this one.  A staging buffer is refilled
    only after the copy out of it has completed."""
    torch = _torch()
    if chunk < 1:
        raise ValueError(f"chunk must be >= 1, got {chunk}")
    mm = np.memmap(path, dtype=np.uint8, mode="r")
    layout = scan_segy(mm)
    dev = torch.device(device) if device is not None else torch.device(
        "cuda", torch.cuda.current_device())
    st = torch.cuda.current_stream(dev)
    body = mm[layout.data_start:]
    end = np.append(layout.trace_pos[1:], body.size) if layout.n_traces else np.zeros(0)
    slots = [None, None]
    copied = [None, None]
    for k, t0 in enumerate(range(0, layout.n_traces, chunk)):
        t1 = min(t0 + chunk, layout.n_traces)
        a, b = int(layout.trace_pos[t0]), int(end[t1 - 1])
        slot = k & 1
        if slots[slot] is None or slots[slot].numel() < b - a:
            if copied[slot] is not None:
                copied[slot].synchronize()
            slots[slot] = torch.empty(b - a, dtype=torch.uint8, pin_memory=True)
        elif copied[slot] is not None:
            copied[slot].synchronize()
        stage = slots[slot][:b - a]
        stage.numpy()[:] = body[a:b]
        dev_bytes = stage.to(dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(st)
        copied[slot] = ev
        sub = SegyLayout(layout.format, layout.sample_bytes, layout.dt_us, layout.n_samples,
                         layout.data_start + a, layout.trace_pos[t0:t1] - a,
                         layout.lengths[t0:t1], layout.fixed)
        res = decode_segy(dev_bytes, sub)
        out = SegyTraces(layout=sub, dt=layout.dt_us * 1e-6)
        if sub.fixed:
            out.signals = res
        else:
            out.values, out.offsets = res
        yield t0, out


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a)


def save_csr(path: str, values, offsets) -> None:
    """A CSR batch (float64 values, int64 offsets[B+1]) as ``.npz``."""
    v, o = _host(values), _host(offsets)
    if v.dtype != np.float64 or o.dtype != np.int64 or o.ndim != 1 or o.size < 1:
        raise ValueError("values must be float64 and offsets int64 [B+1]")
    if o[0] != 0 or o[-1] != v.size or np.any(np.diff(o) < 0):
        raise ValueError("offsets must start at 0, be non-decreasing and end at len(values)")
    np.savez(path, values=v, offsets=o)


def load_csr(path: str):
    """(values, offsets) as NumPy arrays, checked as in save_csr."""
    with np.load(path, allow_pickle=False) as z:
        v, o = z["values"], z["offsets"]
    if v.dtype != np.float64 or o.dtype != np.int64 or o.ndim != 1 or o.size < 1 \
            or o[0] != 0 or o[-1] != v.size or np.any(np.diff(o) < 0):
        raise ValueError(f"{path} is not a CSR batch (float64 values, int64 offsets[B+1])")
    return v, o


def save_compacted(path: str, compact, row_offsets, n_imfs, offsets) -> None:
    """Compacted IMF rows as written by emd_compact: trace s's n_imfs[s] rows of
    len(s) = offsets[s+1] - offsets[s] samples start at compact[row_offsets[s]]."""
    c, r, n, o = _host(compact), _host(row_offsets), _host(n_imfs), _host(offsets)
    if r.size != n.size + 1 or o.size != n.size + 1 or r[-1] > c.size:
        raise ValueError("row_offsets and offsets must have len(n_imfs) + 1 entries")
    np.savez(path, compact=c[:int(r[-1])], row_offsets=r.astype(np.int64),
             n_imfs=n.astype(np.int32), offsets=o.astype(np.int64))


def load_compacted(path: str):
    """(compact, row_offsets, n_imfs, offsets) as NumPy arrays."""
    with np.load(path, allow_pickle=False) as z:
        return z["compact"], z["row_offsets"], z["n_imfs"], z["offsets"]


def trace_rows(compact, row_offsets, n_imfs, offsets, s: int) -> np.ndarray:
    """Trace s's IMF rows [n_imfs[s], len(s)] out of a compacted output (the counterpart
    of GPUBatchEMD.unpack for compacted rows)."""
    n = int(n_imfs[s])
    length = int(offsets[s + 1] - offsets[s])
    a = int(row_offsets[s])
    return np.asarray(compact[a:a + n * length]).reshape(n, length)
