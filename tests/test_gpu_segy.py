"""SEG-Y decode on the GPU: every sample bit-identical to the oracle's decode for each
format (random bytes, so NaNs, infinities, subnormals and both zeros occur), for fixed and
ragged layouts and for traces not aligned to the sample size; and a read SEG-Y batch
decomposes exactly as the same samples handed over as float64."""
import ctypes

import numpy as np
import pytest
import torch

import segy_files
from cuda_emd import GPUBatchEMD, EMDKernelConfig, _lib, synth, traceio
from cuda_emd.kernels import _ptr, _stream
from oracle import segy_ref

pytestmark = pytest.mark.gpu


def _bits(a: np.ndarray) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def _read(tmp_path, data: bytes, name: str):
    p = tmp_path / name
    p.write_bytes(data)
    return traceio.read_segy(str(p))


@pytest.mark.parametrize("fmt", [1, 2, 3, 5, 6, 8])
@pytest.mark.parametrize("N", [2001, 1, 1027])
def test_decode_fixed_bit_exact(tmp_path, fmt, N):
    B = 37
    print(f"segy fixed fmt={fmt} N={N} B={B}", flush=True)
    data, raws = segy_files.random_segy(fmt, [N] * B, seed=100 + fmt + N)
    tr = _read(tmp_path, data, f"f{fmt}_{N}.sgy")
    assert tr.layout.fixed and tr.signals.shape == (B, N) and tr.dt == 0.004
    got = tr.signals.cpu().numpy()
    for b in range(B):
        assert np.array_equal(_bits(got[b]), _bits(segy_ref.decode_samples(raws[b], fmt))), b


@pytest.mark.parametrize("fmt", [1, 3, 5, 6, 8])
def test_decode_ragged_bit_exact(tmp_path, fmt):
    rng = np.random.default_rng(fmt)
    lengths = [int(n) for n in rng.integers(1, 3000, 29)] + [1, 2, 3]
    print(f"segy ragged fmt={fmt} B={len(lengths)}", flush=True)
    data, raws = segy_files.random_segy(fmt, lengths, seed=200 + fmt, ns=lengths[0])
    tr = _read(tmp_path, data, f"r{fmt}.sgy")
    assert not tr.layout.fixed
    v, o = tr.values.cpu().numpy(), tr.offsets.cpu().numpy()
    assert np.array_equal(o, np.concatenate([[0], np.cumsum(lengths)]))
    for b, r in enumerate(raws):
        assert np.array_equal(_bits(v[o[b]:o[b + 1]]), _bits(segy_ref.decode_samples(r, fmt))), b


def test_decode_at_odd_device_address():
    """A byte buffer starting one byte past an aligned address: every trace takes the
    byte-by-byte path, with the same results."""
    fmt, N, B = 1, 333, 9
    data, raws = segy_files.random_segy(fmt, [N] * B, seed=5)
    lay = traceio.scan_segy(np.frombuffer(data, np.uint8))
    body = np.frombuffer(data, np.uint8)[lay.data_start:]
    dev = torch.zeros(body.size + 1, dtype=torch.uint8, device="cuda")
    dev[1:] = torch.from_numpy(body.copy()).cuda()
    out = traceio.decode_segy(dev[1:], lay)
    got = out.cpu().numpy()
    for b in range(B):
        assert np.array_equal(_bits(got[b]), _bits(segy_ref.decode_samples(raws[b], fmt)))


def test_decode_abi_errors():
    lib = _lib.load()
    assert lib.emd_segy_sample_bytes(1) == 4 and lib.emd_segy_sample_bytes(6) == 8
    assert lib.emd_segy_sample_bytes(4) == 0
    out = torch.empty(10, dtype=torch.float64, device="cuda")
    buf = torch.zeros(4096, dtype=torch.uint8, device="cuda")
    rc = lib.emd_segy_decode(_ptr(buf), None, ctypes.c_int64(240), 240, None, 10, 1, 4,
                             _ptr(out), _stream())
    with pytest.raises(ValueError, match="format code 4"):
        _lib.check(rc)
    rc = lib.emd_segy_decode(_ptr(buf), None, ctypes.c_int64(250), 240, None, 10, 1, 1,
                             _ptr(out), _stream())
    with pytest.raises(ValueError, match="shorter than a header"):
        _lib.check(rc)
    assert lib.emd_segy_decode(None, None, ctypes.c_int64(280), 240, None, 10, 0, 1, None,
                               _stream()) == 0


def test_segy_batch_decomposes_like_float_input(tmp_path):
    """IBM-encoded seismic traces read from a file give the same decomposition, bit for
    bit, as the oracle-decoded samples handed to GPUBatchEMD directly."""
    B, N = 48, 2001
    x = synth.seismic_traces(B, N, seed=11)
    raws = [segy_files.encode_samples(x[b] * 1000.0, 1) for b in range(B)]
    data = segy_files.segy_bytes(raws, [N] * B, 1)
    print(f"segy end-to-end B={B} N={N}", flush=True)
    tr = _read(tmp_path, data, "seis.sgy")
    ref = np.stack([segy_ref.decode_samples(r, 1) for r in raws])
    assert np.array_equal(_bits(tr.signals.cpu().numpy()), _bits(ref))
    emd = GPUBatchEMD(EMDKernelConfig.for_device())
    a = emd(tr.signals)
    da = emd.last_debug
    b = emd(torch.from_numpy(ref).cuda())
    db = emd.last_debug
    for u, v in zip(list(a) + list(da), list(b) + list(db)):
        assert torch.equal(u, v)


@pytest.mark.parametrize("fmt,ragged", [(1, False), (3, True), (5, False)])
def test_iter_segy_chunks_match_whole_read(tmp_path, fmt, ragged):
    rng = np.random.default_rng(30 + fmt)
    lengths = [int(n) for n in rng.integers(1, 900, 37)] if ragged else [517] * 37
    print(f"iter_segy fmt={fmt} ragged={ragged}", flush=True)
    data, _ = segy_files.random_segy(fmt, lengths, seed=40 + fmt, ns=lengths[0])
    p = tmp_path / "it.sgy"
    p.write_bytes(data)
    whole = traceio.read_segy(str(p))
    seen = 0
    for t0, part in traceio.iter_segy(str(p), chunk=10):
        assert t0 == seen
        n = part.layout.n_traces
        if ragged:
            o = whole.offsets.cpu().numpy()
            got = part.values.cpu().numpy()
            assert np.array_equal(part.offsets.cpu().numpy(), o[t0:t0 + n + 1] - o[t0])
            assert np.array_equal(_bits(got), _bits(whole.values.cpu().numpy()[o[t0]:o[t0 + n]]))
        else:
            assert np.array_equal(_bits(part.signals.cpu().numpy()),
                                  _bits(whole.signals.cpu().numpy()[t0:t0 + n]))
        seen += n
    assert seen == len(lengths)


def test_read_segy_without_traces(tmp_path):
    """A file holding only its 3600-byte header reads as an empty [0, ns] batch."""
    data = segy_files.segy_bytes([], [], 1, ns=250)
    print("segy header only", flush=True)
    tr = _read(tmp_path, data, "empty.sgy")
    assert tr.layout.fixed and tuple(tr.signals.shape) == (0, 250)
    assert list(traceio.iter_segy(str(tmp_path / "empty.sgy"))) == []
