"""The Hilbert spectral step (emd_hilbert through cuda_emd.hilbert) against the oracle
(oracle/hilbert_ref.py: scipy.signal.hilbert, numpy.angle / unwrap).

Tolerances, per row with s = max|x|: amplitude within 1e-12 s; phase within 1e-9 rad
(wrapped) and frequency within 1e-9 cycles per sample (wrapped, i.e. times dt) wherever the
oracle's amplitude is at least 1e-3 of the row's largest (elsewhere the phase is
ill-conditioned: a rounding-level change of z turns it by up to ~1e-13 s / |z|)."""
import numpy as np
import pytest

from oracle import hilbert_ref as H

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def _check_row(x, amp, phase, freq, dt, n_fft, label):
    ra, rp, rf = H.instantaneous(x, dt=dt, n_fft=n_fft)
    s = float(np.max(np.abs(x))) if len(x) else 0.0
    if s == 0.0:
        for got in (amp, phase, freq):
            if got is not None:
                assert not np.any(got), label
        return
    if amp is not None:
        assert np.max(np.abs(amp - ra)) <= 1e-12 * s, (label, np.max(np.abs(amp - ra)) / s)
    well = ra >= 1e-3 * np.max(ra)
    if phase is not None:
        d = np.angle(np.exp(1j * (phase - rp)))
        assert np.max(np.abs(d[well]), initial=0.0) <= 1e-9, label
    if freq is not None and len(x) > 1:
        n = np.arange(len(x))
        m = np.minimum(n, len(x) - 2)
        ok = well[m] & well[m + 1]
        d = (freq - rf) * dt
        d = d - np.round(d)
        assert np.max(np.abs(d[ok]), initial=0.0) <= 1e-9, label
    elif freq is not None:
        assert freq[0] == 0.0, label


def _padded_rows(B, K, N, seed):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((B, K, N))
    x[:, 1] = np.cumsum(x[:, 1], axis=-1)  # a red row beside the white ones
    return x


@pytest.mark.parametrize("N", [7, 1000, 1024, 2001])
@pytest.mark.parametrize("n_fft", [0, -1, 4096])
def test_padded(N, n_fft):
    torch = _torch()
    from cuda_emd.hilbert import hilbert_spectrum
    B, K, dt = 6, 5, 0.004
    x = _padded_rows(B, K, N, seed=N + 7 * n_fft)
    n_imfs = np.array([5, 4, 3, 1, 0, 2], np.int32)
    print(f"hilbert padded N={N} n_fft={n_fft}", flush=True)
    out = hilbert_spectrum(torch.from_numpy(x).cuda(), torch.from_numpy(n_imfs).cuda(), dt=dt,
                           n_fft=n_fft)
    torch.cuda.synchronize()
    amp, phase, freq = (out[k].cpu().numpy() for k in ("amp", "phase", "freq"))
    for b in range(B):
        for k in range(K):
            if k < n_imfs[b]:
                _check_row(x[b, k], amp[b, k], phase[b, k], freq[b, k], dt, n_fft,
                           (N, n_fft, b, k))
            else:
                assert not (amp[b, k].any() or phase[b, k].any() or freq[b, k].any())


def _ragged(lengths, seed, max_rows=4):
    rng = np.random.default_rng(seed)
    lengths = np.asarray(lengths, np.int64)
    B = len(lengths)
    n_imfs = rng.integers(0, max_rows + 1, B).astype(np.int32)
    n_imfs[:min(B, 3)] = [3, 1, max_rows][:min(B, 3)]
    offsets = np.concatenate([[0], np.cumsum(lengths)])
    row_offsets = np.concatenate([[0], np.cumsum(n_imfs * lengths)])
    flat = rng.standard_normal(int(row_offsets[-1]))
    return flat, offsets, row_offsets, n_imfs


LENGTHS = [1, 2, 3, 4, 5, 8, 17, 64, 255, 256, 257, 1000, 2001, 4096, 4097, 8191, 8192]


@pytest.mark.parametrize("n_fft", [0, -1, 16384])
def test_ragged_compact_layout(n_fft):
    """CSR traces of every transform class in one batch, rows packed back to back
    (row_base = the compaction's row offsets)."""
    torch = _torch()
    from cuda_emd.hilbert import hilbert_spectrum
    rng = np.random.default_rng(5)
    lengths = LENGTHS + list(rng.integers(1, 3000, 40))
    flat, offsets, row_offsets, n_imfs = _ragged(lengths, seed=9 + n_fft)
    dt = 0.002
    print(f"hilbert ragged compact n_fft={n_fft} B={len(lengths)}", flush=True)
    out = hilbert_spectrum(torch.from_numpy(flat).cuda(), torch.from_numpy(n_imfs).cuda(),
                           offsets=torch.from_numpy(offsets).cuda(),
                           row_base=torch.from_numpy(row_offsets[:-1]).cuda(), dt=dt,
                           n_fft=n_fft)
    torch.cuda.synchronize()
    amp, phase, freq = (out[k].cpu().numpy() for k in ("amp", "phase", "freq"))
    for s, L in enumerate(lengths):
        for k in range(n_imfs[s]):
            a = row_offsets[s] + k * L
            sl = slice(a, a + L)
            _check_row(flat[sl], amp[sl], phase[sl], freq[sl], dt, n_fft, (n_fft, s, L, k))


def test_ragged_default_layout_and_many_traces():
    """run_ragged's default layout (trace s's rows from offsets[s] * max_imfs) with more
    traces than the resident workgroups, so every class launch walks the batch."""
    torch = _torch()
    from cuda_emd.hilbert import hilbert_spectrum
    rng = np.random.default_rng(21)
    B, K = 6000, 3
    lengths = rng.integers(1, 700, B)
    lengths[::997] = 2001
    offsets = np.concatenate([[0], np.cumsum(lengths)])
    n_imfs = rng.integers(0, K + 1, B).astype(np.int32)
    flat = rng.standard_normal(int(offsets[-1]) * K)
    print(f"hilbert ragged default B={B}", flush=True)
    out = hilbert_spectrum(torch.from_numpy(flat).cuda(), torch.from_numpy(n_imfs).cuda(),
                           offsets=torch.from_numpy(offsets).cuda(), max_imfs=K,
                           outputs=("amp", "freq"))
    torch.cuda.synchronize()
    assert set(out) == {"amp", "freq"}
    amp, freq = out["amp"].cpu().numpy(), out["freq"].cpu().numpy()
    for s in list(range(0, B, 7)) + list(range(0, B, 997)):
        L = lengths[s]
        for k in range(n_imfs[s]):
            a = offsets[s] * K + k * L
            sl = slice(a, a + L)
            _check_row(flat[sl], amp[sl], None, freq[sl], 1.0, 0, (s, L, k))


def test_emd_output_padded_and_compact_agree():
    """The step on a real decomposition: GPUBatchEMD's padded IMFs and their compaction give
    bit-identical spectra (same kernel, same rows), and both match the oracle."""
    torch = _torch()
    from cuda_emd import GPUBatchEMD, synth
    from cuda_emd.hilbert import hilbert_spectrum
    x = synth.seismic_traces(48, 2001, seed=1)
    emd = GPUBatchEMD()
    imfs, n_imfs, _ = emd(torch.from_numpy(x).cuda())
    print("hilbert on seismic IMFs", flush=True)
    pad = hilbert_spectrum(imfs, n_imfs, dt=0.004)
    row, comp = emd._get_launcher().compact(imfs, n_imfs)
    offsets = torch.arange(0, 49 * 2001, 2001, dtype=torch.int64, device="cuda")
    com = hilbert_spectrum(comp, n_imfs, offsets=offsets, row_base=row[:-1], dt=0.004)
    torch.cuda.synchronize()
    nim = n_imfs.cpu().numpy()
    rows = row.cpu().numpy()
    imfs_h = imfs.cpu().numpy()
    for key in ("amp", "phase", "freq"):
        p, c = pad[key].cpu().numpy(), com[key].cpu().numpy()
        for s in range(len(x)):
            got = c[rows[s]:rows[s + 1]].reshape(nim[s], 2001)
            assert np.array_equal(got, p[s, :nim[s]]), (key, s)
    amp, phase, freq = (pad[k].cpu().numpy() for k in ("amp", "phase", "freq"))
    for s in range(0, len(x), 5):
        for k in range(nim[s]):
            _check_row(imfs_h[s, k], amp[s, k], phase[s, k], freq[s, k], 0.004, 0, (s, k))


def test_batch_composition_invariance():
    """A trace's spectra do not depend on the rest of the batch (the class launches skip
    other classes' traces and a trace's rows are transformed by one workgroup)."""
    torch = _torch()
    from cuda_emd.hilbert import hilbert_spectrum
    lengths = [2001, 300, 4097, 2001, 16, 1024]
    flat, offsets, row_offsets, n_imfs = _ragged(lengths, seed=33, max_rows=5)
    kw = dict(dt=0.004)
    print("hilbert batch invariance", flush=True)
    full = hilbert_spectrum(torch.from_numpy(flat).cuda(), torch.from_numpy(n_imfs).cuda(),
                            offsets=torch.from_numpy(offsets).cuda(),
                            row_base=torch.from_numpy(row_offsets[:-1]).cuda(), **kw)
    for s in (0, 2, 5):
        a, b = row_offsets[s], row_offsets[s + 1]
        one = hilbert_spectrum(torch.from_numpy(flat[a:b].copy()).cuda(),
                               torch.tensor([n_imfs[s]], dtype=torch.int32).cuda(),
                               offsets=torch.tensor([0, lengths[s]], dtype=torch.int64).cuda(),
                               row_base=torch.zeros(1, dtype=torch.int64).cuda(), **kw)
        for key in one:
            assert np.array_equal(one[key].cpu().numpy(), full[key][a:b].cpu().numpy()), (s, key)
