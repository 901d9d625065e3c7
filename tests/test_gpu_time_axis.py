"""The time-axis kernels (EMDKernelConfig(time_axis='given')): the reference's use of T,
knots at T[idx], mirrored knots at 2 T[sym] - T[idx] and envelopes at T[i]
(kernels.py:471-472, 528-552, 703), after the reference's normalisation (emd.py:83-88),
against the oracle given the same T, through the C ABI (emd_run_padded_times,
emd_run_ragged_times, emd_envelope_only_times)."""
import numpy as np
import pytest

from oracle import pyemd_restated as O
from parity import compare

pytestmark = pytest.mark.gpu


def _jittered(rng, n, lo=0.6, hi=1.4, t0=0.0, dt=0.004):
    return t0 + dt * np.cumsum(rng.uniform(lo, hi, n))


def _records(imfs, n, ovf, dn, de, dz, stats):
    return dict(imfs=imfs.cpu().numpy(), n_imfs=n.cpu().numpy(), overflow=ovf.cpu().numpy(),
                debug_n=dn.cpu().numpy(), debug_extno=de.cpu().numpy(),
                debug_nzm=dz.cpu().numpy(), stats=stats.cpu().numpy())


def _gpu_emd(signals, T, **cfg):
    import torch
    from cuda_emd import EMDKernelConfig, GPUBatchEMD
    g = GPUBatchEMD(EMDKernelConfig(time_axis="given", **cfg))
    imfs, n, ovf = g(torch.from_numpy(np.ascontiguousarray(signals)).cuda(), T)
    torch.cuda.synchronize()
    return _records(imfs, n, ovf, *g.last_debug, g.last_stats)


def _assert_parity(signals, gpu, ora, label):
    bad, worst = compare(list(signals), gpu, ora, 16, label=label)
    unexplained = [m for m in bad if not m[2]]
    print(f"{label}: {len(signals)} traces, {len(bad)} mismatches ({len(unexplained)} "
          f"unexplained), worst IMF err {worst:.3e} x max|x|")
    assert not unexplained, unexplained
    return bad


def test_shared_jittered_axis():
    """One jittered axis for the whole batch (the reference's [N] layout)."""
    from cuda_emd import synth
    rng = np.random.default_rng(21)
    x = np.concatenate([synth.chirp_mix(24, 1024, seed=3), synth.seismic_traces(8, 1024, seed=4)])
    T = _jittered(rng, x.shape[1])
    ora = [O.emd(s, max_imf=15, T=T) for s in x]
    _assert_parity(x, _gpu_emd(x, T), ora, "time-axis shared")


def test_per_signal_axes():
    """One axis per signal [B,N], each normalised on its own; a T that differs from the
    sample index must change the decomposition."""
    from cuda_emd import synth
    rng = np.random.default_rng(22)
    x = synth.chirp_mix(32, 800, seed=6)
    T = np.stack([_jittered(rng, 800, 0.5, 1.5, t0=float(k)) for k in range(32)])
    ora = [O.emd(s, max_imf=15, T=t) for s, t in zip(x, T)]
    gpu = _gpu_emd(x, T)
    _assert_parity(x, gpu, ora, "time-axis per-signal")
    plain = [O.emd(s, max_imf=15) for s in x]
    assert sum(int(np.array_equal(a.debug_n, b.debug_n)) for a, b in zip(ora, plain)) < len(x)


def test_index_axis_through_time_kernels_is_identical():
    """T whose normalisation is exactly the sample index (0.5 n + 3) runs the time-axis
    kernels with every position an exact integer: the outputs equal the index kernels' bit
    for bit."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    x = synth.seismic_traces(16, 2001, seed=1)
    T = 0.5 * np.arange(2001) + 3.0
    timed = _gpu_emd(x, T)
    L = EMDKernelConfig().build()
    ref = _records(*L.run(torch.from_numpy(x).cuda()), L.last_stats)
    for k in ("n_imfs", "overflow", "debug_n", "debug_extno", "debug_nzm", "imfs"):
        assert np.array_equal(timed[k], ref[k]), k
    assert np.array_equal(timed["stats"][:, :3], ref["stats"][:, :3])


def test_ragged_times_persistent_small_tile():
    """CSR traces of uneven length with per-trace axes, persistent at 64 lanes with a
    60-slot LDS tile (the knots go to the global tile of 32-byte slots), against the oracle
    and against the plain 256-lane launch bit for bit."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    rng = np.random.default_rng(23)
    base = synth.chirp_mix(20, 1500, seed=8)
    traces = [np.ascontiguousarray(base[k, :int(rng.integers(300, 1501))]) for k in range(20)]
    Ts = [_jittered(rng, len(t)) for t in traces]
    off = np.concatenate([[0], np.cumsum([len(t) for t in traces])]).astype(np.int64)
    vals = torch.from_numpy(np.concatenate(traces)).cuda()
    times = torch.from_numpy(np.concatenate([O.normalize_time(t) for t in Ts])).cuda()
    L = EMDKernelConfig().build()

    def run(**kw):
        imfs, oo, n, ovf, dn, de, dz, st = L.run_ragged(vals, off, times=times, **kw)
        torch.cuda.synchronize()
        oo, n = oo.cpu().numpy(), n.cpu().numpy()
        flat = imfs.cpu().numpy()
        rows = [flat[oo[s]: oo[s] + n[s] * len(traces[s])].reshape(n[s], len(traces[s]))
                for s in range(len(traces))]
        return dict(imfs=rows, n_imfs=n, overflow=ovf.cpu().numpy(), debug_n=dn.cpu().numpy(),
                    debug_extno=de.cpu().numpy(), debug_nzm=dz.cpu().numpy(),
                    stats=st.cpu().numpy())

    small = run(threads=64, tile_slots=60, persistent=True)
    plain = run(threads=256, persistent=False)
    ora = [O.emd(s, max_imf=15, T=t) for s, t in zip(traces, Ts)]
    _assert_parity(traces, small, ora, "time-axis ragged")
    for k in ("n_imfs", "overflow", "debug_n", "debug_extno", "debug_nzm"):
        assert np.array_equal(small[k], plain[k]), k
    for a, b in zip(small["imfs"], plain["imfs"]):
        assert np.array_equal(a, b)


def test_envelope_only_with_axis():
    """Envelopes and knot positions (in time units) of one pass, against the oracle."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    rng = np.random.default_rng(24)
    x = synth.chirp_mix(16, 700, seed=9)
    T = _jittered(rng, 700)
    L = EMDKernelConfig(time_axis="given").build()
    Tn = O.normalize_time(T)
    out = L.run_envelope_only(torch.from_numpy(x).cuda(), torch.from_numpy(Tn).cuda())
    torch.cuda.synchronize()
    mx, mn, n_ext, deg, mkp, mkv, nkp, nkv, n_mir = (t.cpu().numpy() for t in out)
    for b in range(len(x)):
        r = O.envelopes(x[b], T=T)
        assert deg[b] == r["degenerate"] == 0
        scale = np.max(np.abs(x[b]))
        assert np.max(np.abs(mx[b] - r["max_env"])) <= 1e-9 * scale
        assert np.max(np.abs(mn[b] - r["min_env"])) <= 1e-9 * scale
        k = n_mir[b]
        assert k[0] == r["max_knots"].shape[1] and k[1] == r["min_knots"].shape[1]
        assert np.array_equal(mkp[b, :k[0]], r["max_knots"][0])
        assert np.array_equal(nkp[b, :k[1]], r["min_knots"][0])
        assert np.array_equal(mkv[b, :k[0]], r["max_knots"][1])


def test_edge_signals_and_short_lengths_on_an_axis():
    """The index-axis edge signals (plateaus, exact zeros, steps, spikes, square waves)
    and lengths 0-130 on jittered axes: the plateau rules, the small spline kinds and the
    degenerate traces with time positions, against the oracle."""
    from test_gpu_parity import _edge_signals
    rng = np.random.default_rng(25)
    sigs = list(_edge_signals().values())
    T = _jittered(rng, sigs[0].size)
    ora = [O.emd(s, max_imf=15, T=T) for s in sigs]
    x = np.stack(sigs)
    _assert_parity(x, _gpu_emd(x, T), ora, "time-axis edge signals")
    for N in (0, 1, 2, 3, 4, 5, 7, 64, 65, 130):
        x = rng.standard_normal((3, N))
        T = np.stack([_jittered(rng, N) for _ in range(3)]) if N else np.zeros((3, 0))
        ora = [O.emd(s, max_imf=15, T=t) for s, t in zip(x, T)]
        _assert_parity(x, _gpu_emd(x, T), ora, f"time-axis N={N}")
