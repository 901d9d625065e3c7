"""The SEG-Y decode oracle pinned against textbook IBM values and exact rational arithmetic,
and the host side of the reader (header parsing, trace layout, CSR / compacted files)."""
from fractions import Fraction

import numpy as np
import pytest

from cuda_emd import traceio
from oracle import segy_ref
import segy_files


def _ibm_exact(w: int) -> Fraction:
    sign = -1 if w >> 31 else 1
    e = (w >> 24) & 0x7F
    f = w & 0xFFFFFF
    return sign * Fraction(f, 1 << 24) * Fraction(16) ** (e - 64)


@pytest.mark.parametrize("word,value", [
    (0xC276A000, -118.625),            # the standard worked example
    (0x42640000, 100.0),
    (0x41100000, 1.0),
    (0xC1100000, -1.0),
    (0x40800000, 0.5),
    (0x3F800000, 0.03125),
    (0x00000000, 0.0),
    (0x7FFFFFFF, float(Fraction(0xFFFFFF, 1 << 24) * Fraction(16) ** 63)),
    (0x00100000, 16.0 ** -65),          # smallest normalised magnitude
    (0x00000001, 2.0 ** -280),          # smallest non-zero (unnormalised) magnitude
])
def test_ibm_known_words(word, value):
    got = segy_ref.ibm32_to_f64(np.array([word], np.uint64))[0]
    assert got == value
    assert np.signbit(got) == (word >> 31 == 1 and value != 0 or word == 0x80000000)


def test_ibm_negative_zero():
    got = segy_ref.ibm32_to_f64(np.array([0x80000000], np.uint64))[0]
    assert got == 0.0 and np.signbit(got)


def test_ibm_random_words_exact():
    rng = np.random.default_rng(7)
    words = rng.integers(0, 2 ** 32, 4000, dtype=np.uint64)
    got = segy_ref.ibm32_to_f64(words)
    for w, g in zip(words.tolist(), got.tolist()):
        assert Fraction(g) == _ibm_exact(w), hex(w)


def test_ibm_encoder_round_trip():
    rng = np.random.default_rng(8)
    x = rng.standard_normal(5000) * 10.0 ** rng.integers(-30, 30, 5000)
    w = segy_files.f64_to_ibm32(x)
    y = segy_ref.ibm32_to_f64(w.astype(np.uint64))
    # Truncation of a 24-bit base-16 fraction: relative error below 16^-5 (leading digit >= 1).
    assert np.all(np.abs(y) <= np.abs(x)) and np.all(np.sign(y) == np.sign(x))
    assert np.max(np.abs(y - x) / np.abs(x)) < 16.0 ** -5
    # Values the format holds exactly come back unchanged.
    assert np.array_equal(segy_ref.ibm32_to_f64(segy_files.f64_to_ibm32(y).astype(np.uint64)), y)


@pytest.mark.parametrize("fmt,raw,value", [
    (5, b"\x3f\x80\x00\x00", 1.0), (5, b"\xc0\x49\x0f\xdb", float(np.float32(-3.14159265))),
    (2, b"\xff\xff\xff\xfe", -2.0), (3, b"\x80\x00", -32768.0), (8, b"\x7f", 127.0),
    (6, b"\x3f\xf0\x00\x00\x00\x00\x00\x00", 1.0), (1, b"\xc2\x76\xa0\x00", -118.625),
])
def test_decode_samples_known(fmt, raw, value):
    assert segy_ref.decode_samples(raw, fmt)[0] == value


def _layout_of(data: bytes):
    return traceio.scan_segy(np.frombuffer(data, np.uint8))


def test_scan_fixed_length():
    data, raws = segy_files.random_segy(1, [501] * 7, seed=1, n_ext=2)
    lay = _layout_of(data)
    assert lay.fixed and lay.n_traces == 7 and lay.n_samples == 501 and lay.format == 1
    assert lay.data_start == 3600 + 2 * 3200 and lay.dt_us == 4000
    assert np.array_equal(lay.trace_pos, np.arange(7) * (240 + 4 * 501))
    got = segy_ref.decode_traces(np.frombuffer(data, np.uint8)[lay.data_start:], lay.trace_pos,
                                 lay.lengths, 240, 1)
    for g, r in zip(got, raws):
        assert np.array_equal(g.view(np.uint64),
                              segy_ref.decode_samples(r, 1).view(np.uint64))


def test_scan_zero_trace_sample_counts_use_binary_header():
    raws = [bytes(2 * 33)] * 4
    data = segy_files.segy_bytes(raws, [33] * 4, 3, own_ns=False)
    lay = _layout_of(data)
    assert lay.fixed and np.array_equal(lay.lengths, [33] * 4)


def test_scan_ragged():
    lengths = [300, 17, 1, 1201, 5]
    data, _ = segy_files.random_segy(3, lengths, seed=2, ns=300)
    lay = _layout_of(data)
    assert not lay.fixed and np.array_equal(lay.lengths, lengths)
    assert np.array_equal(lay.offsets, np.concatenate([[0], np.cumsum(lengths)]))
    pos = np.concatenate([[0], np.cumsum(240 + 2 * np.array(lengths))[:-1]])
    assert np.array_equal(lay.trace_pos, pos)


def test_scan_errors():
    data, _ = segy_files.random_segy(1, [100] * 3, seed=3)
    bad = bytearray(data)
    bad[3224:3226] = b"\x00\x04"                # format 4 (fixed point with gain): unsupported
    with pytest.raises(ValueError, match="format code 4"):
        _layout_of(bytes(bad))
    with pytest.raises(ValueError, match="past the end"):
        _layout_of(data[:-7])
    with pytest.raises(ValueError, match="3600-byte"):
        _layout_of(data[:1000])


def test_csr_and_compacted_files(tmp_path):
    rng = np.random.default_rng(4)
    offsets = np.array([0, 5, 5, 12], np.int64)
    values = rng.standard_normal(12)
    p = str(tmp_path / "batch.npz")
    traceio.save_csr(p, values, offsets)
    v, o = traceio.load_csr(p)
    assert np.array_equal(v, values) and np.array_equal(o, offsets)
    with pytest.raises(ValueError):
        traceio.save_csr(p, values, np.array([0, 5, 4, 12], np.int64))
    n_imfs = np.array([2, 1, 3], np.int32)
    lens = np.diff(offsets)
    rows = np.concatenate([[0], np.cumsum(n_imfs * lens)]).astype(np.int64)
    compact = rng.standard_normal(int(rows[-1]))
    q = str(tmp_path / "rows.npz")
    traceio.save_compacted(q, compact, rows, n_imfs, offsets)
    c, r, n, o = traceio.load_compacted(q)
    assert np.array_equal(c, compact) and np.array_equal(r, rows)
    t2 = traceio.trace_rows(c, r, n, o, 2)
    assert t2.shape == (3, 7) and np.array_equal(t2.ravel(), compact[rows[2]:rows[3]])
    assert traceio.trace_rows(c, r, n, o, 1).shape == (1, 0)
