"""Builds SEG-Y files for the tests: a 3200-byte textual header, a 400-byte binary header
(sample interval, samples per trace, format code, extended header count), optional
extended textual headers, then 240-byte trace headers each followed by its samples.
Positions are the SEG-Y rev 1 standard's (1-based byte numbers in comments)."""
from __future__ import annotations

import struct

import numpy as np

SAMPLE_BYTES = {1: 4, 2: 4, 3: 2, 5: 4, 6: 8, 8: 1}


def f64_to_ibm32(x: np.ndarray) -> np.ndarray:
    """IBM single words nearest below |x| in magnitude (fraction truncated), sign kept.
    |x| = m * 2^p (m in [0.5, 1)) is written as 16^E * (m * 2^(p - 4E)) with E = ceil(p / 4),
    so the base-16 fraction lies in [1/16, 1)."""
    x = np.asarray(x, dtype=np.float64)
    m, p = np.frexp(np.abs(x))
    E = -((-p) // 4)
    frac = np.floor(np.ldexp(m, (p - 4 * E + 24).astype(np.int32))).astype(np.uint64)
    e = (E + 64).astype(np.int64)
    if np.any((x != 0) & ((e < 0) | (e > 127))):
        raise ValueError("value outside the IBM single range")
    w = (np.where(x != 0, e, 0).astype(np.uint64) << np.uint64(24)) | np.where(x != 0, frac, 0)
    return (w | (np.signbit(x).astype(np.uint64) << np.uint64(31))).astype(np.uint32)


def encode_samples(x: np.ndarray, fmt: int) -> bytes:
    """Float samples as big-endian bytes of the given format (integers are rounded)."""
    x = np.asarray(x, dtype=np.float64)
    if fmt == 1:
        return f64_to_ibm32(x).astype(">u4").tobytes()
    dt = {2: ">i4", 3: ">i2", 5: ">f4", 6: ">f8", 8: "i1"}[fmt]
    if fmt in (2, 3, 8):
        x = np.rint(x)
    return x.astype(dt).tobytes()


def binary_header(ns: int, fmt: int, dt_us: int = 4000, n_ext: int = 0) -> bytes:
    h = bytearray(400)
    struct.pack_into(">h", h, 16, dt_us)      # 3217-3218 sample interval (us)
    struct.pack_into(">H", h, 20, ns)         # 3221-3222 samples per trace
    struct.pack_into(">h", h, 24, fmt)        # 3225-3226 sample format code
    struct.pack_into(">H", h, 300, 0x0100)    # 3501-3502 revision 1.0
    struct.pack_into(">h", h, 302, 1)         # 3503-3504 fixed-length traces
    struct.pack_into(">h", h, 304, n_ext)     # 3505-3506 extended textual headers
    return bytes(h)


def trace_header(seq: int, ns: int, dt_us: int = 4000) -> bytes:
    h = bytearray(240)
    struct.pack_into(">i", h, 0, seq)         # 1-4 trace sequence number
    struct.pack_into(">H", h, 114, ns)        # 115-116 samples in this trace
    struct.pack_into(">H", h, 116, dt_us)     # 117-118 sample interval (us)
    return bytes(h)


def segy_bytes(samples: list[bytes], lengths: list[int], fmt: int, ns: int | None = None,
               dt_us: int = 4000, n_ext: int = 0, own_ns: bool = True) -> bytes:
    """A whole file from each trace's encoded samples.  ``own_ns`` False writes 0 in the
    trace headers' sample counts (readers then use the binary header's)."""
    ns = lengths[0] if ns is None and lengths else (ns or 0)
    parts = [b"C" * 3200, binary_header(ns, fmt, dt_us, n_ext), b" " * (3200 * n_ext)]
    for i, (raw, n) in enumerate(zip(samples, lengths)):
        parts.append(trace_header(i + 1, n if own_ns else 0, dt_us))
        parts.append(raw)
    return b"".join(parts)


def random_segy(fmt: int, lengths: list[int], seed: int, n_ext: int = 0, ns: int | None = None):
    """A file whose samples are random bytes (every bit pattern of the format, NaNs,
    infinities and subnormals included).  Returns (file bytes, per-trace raw sample bytes)."""
    rng = np.random.default_rng(seed)
    s = SAMPLE_BYTES[fmt]
    raws = [rng.integers(0, 256, n * s, dtype=np.uint8).tobytes() for n in lengths]
    return segy_bytes(raws, lengths, fmt, ns=ns, n_ext=n_ext), raws
