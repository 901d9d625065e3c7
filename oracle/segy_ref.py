"""CPU restatement (NumPy) of SEG-Y sample decoding: big-endian samples of the rev 1 formats
1 (IBM hexadecimal float), 2 (int32), 3 (int16), 5 (IEEE float), 8 (int8) and rev 2's
6 (IEEE double), to float64.

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the HIP decode kernel in
emd-cuda_amd/csrc/emd_segy.hip; the product path never imports it.

The reference has no SEG-Y reader (it takes float64 device arrays, emd.py:60-69, and
describes its inputs as seismic traces at 4 ms, emd.py:28-30), so nothing here follows a
reference file; the formats follow the published SEG-Y rev 1 standard.  Pinning
(tests/test_segy_oracle.py): IBM words against textbook values (0xC276A000 = -118.625,
0x42640000 = 100, 0x41100000 = 1, the largest and smallest magnitudes, both zeros) and,
over random words, against exact rational arithmetic (fraction / 2^24 * 16^(e-64) with
Python's Fraction); the other formats are NumPy's own big-endian dtype conversions.
"""
from __future__ import annotations

import numpy as np

SAMPLE_BYTES = {1: 4, 2: 4, 3: 2, 5: 4, 6: 8, 8: 1}
_BE_DTYPE = {2: ">i4", 3: ">i2", 5: ">f4", 6: ">f8", 8: "i1"}


def ibm32_to_f64(words: np.ndarray) -> np.ndarray:
    """IBM single words (as unsigned ints, most significant byte first) to float64:
    (-1)^s * f / 2^24 * 16^(e - 64), f the 24-bit fraction and e the 7-bit exponent."""
    w = np.asarray(words, dtype=np.uint64)
    frac = (w & 0xFFFFFF).astype(np.float64)
    e = ((w >> 24) & 0x7F).astype(np.int64)
    mag = np.ldexp(frac, (4 * e - 280).astype(np.int32))
    return np.where((w >> 31) != 0, -mag, mag)


def decode_samples(raw: bytes | np.ndarray, fmt: int) -> np.ndarray:
    """Decode a run of big-endian samples (bytes in file order) to float64."""
    buf = np.frombuffer(bytes(raw) if not isinstance(raw, np.ndarray) else raw.tobytes(),
                        dtype=np.uint8)
    if fmt == 1:
        return ibm32_to_f64(buf.view(">u4").astype(np.uint64))
    if fmt not in _BE_DTYPE:
        raise ValueError(f"unsupported SEG-Y format {fmt}")
    with np.errstate(invalid="ignore"):     # NaN samples convert to NaN, as on the device
        return buf.view(_BE_DTYPE[fmt]).astype(np.float64)


def decode_traces(data: np.ndarray, trace_pos: np.ndarray, lengths: np.ndarray,
                  header_bytes: int, fmt: int) -> list[np.ndarray]:
    """Each trace's samples: lengths[b] samples starting header_bytes after trace_pos[b]."""
    s = SAMPLE_BYTES[fmt]
    out = []
    for p, n in zip(np.asarray(trace_pos, np.int64), np.asarray(lengths, np.int64)):
        a = int(p) + header_bytes
        out.append(decode_samples(data[a:a + int(n) * s], fmt))
    return out
