"""HIP path vs the CPU oracle (PyEMD EMD() defaults restated), through the C ABI."""
import numpy as np
import pytest

from oracle import pyemd_restated as O
from parity import compare

pytestmark = pytest.mark.gpu


def _gpu_padded(signals, cfg=None):
    import torch
    from cuda_emd import EMDKernelConfig
    cfg = cfg or EMDKernelConfig()
    launcher = cfg.build()
    x = torch.from_numpy(np.ascontiguousarray(signals)).cuda()
    imfs, n, ovf, dn, de, dz = launcher.run(x)
    torch.cuda.synchronize()
    return dict(imfs=imfs.cpu().numpy(), n_imfs=n.cpu().numpy(), overflow=ovf.cpu().numpy(),
                debug_n=dn.cpu().numpy(), debug_extno=de.cpu().numpy(),
                debug_nzm=dz.cpu().numpy(), stats=launcher.last_stats.cpu().numpy())


def _oracle(signals, max_extrema=None, max_iteration=1000, max_imfs=16):
    return [O.emd(s, max_imf=max_imfs - 1, max_iteration=max_iteration,
                  max_extrema=max_extrema, debug_slots=max_imfs) for s in signals]


def _assert_parity(signals, gpu, ora, max_imfs=16):
    import os
    label = os.environ.get("PYTEST_CURRENT_TEST", "").split(" ")[0]
    bad, worst = compare(list(signals), gpu, ora, max_imfs, label=label)
    unexplained = [m for m in bad if not m[2]]
    print(f"parity: {len(signals)} traces, {len(bad)} mismatches "
          f"({len(unexplained)} unexplained), worst IMF err {worst:.3e} x max|x|")
    for m in bad:
        print("  mismatch", m)
    assert not unexplained, unexplained


def test_chirp_mix_parity():
    from cuda_emd import synth
    x = synth.chirp_mix()[:48]
    _assert_parity(x, _gpu_padded(x), _oracle(x))


def test_seismic_parity():
    from cuda_emd import synth
    x = synth.seismic_traces(24, 2001, seed=1)
    _assert_parity(x, _gpu_padded(x), _oracle(x))


def _edge_signals(N=257):
    n = np.arange(N, dtype=np.float64)
    sigs = {
        "zeros": np.zeros(N),
        "const": np.full(N, 3.5),
        "ramp": 0.01 * n - 1.0,
        "tone": np.sin(2 * np.pi * n / 32.0),
        "tone_trend": np.sin(2 * np.pi * n / 20.0) + 0.002 * n,
        "stairs": np.floor(np.sin(2 * np.pi * n / 50.0) * 4.0),
        "gated": np.where((n // 40) % 2 == 0, np.sin(2 * np.pi * 0.05 * n), 0.0),
        "spikes": np.where(n % 17 == 0, 1.0, 0.0),
        "square": np.sign(np.sin(2 * np.pi * n / 30.0)),
    }
    return sigs


@pytest.mark.parametrize("name", list(_edge_signals().keys()))
def test_edge_signal_parity(name):
    x = _edge_signals()[name][None, :]
    _assert_parity(x, _gpu_padded(x), _oracle(x))


@pytest.mark.parametrize("N", [0, 1, 2, 3, 4, 5, 7, 64, 65, 130])
def test_short_lengths(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((3, N))
    _assert_parity(x, _gpu_padded(x), _oracle(x))


@pytest.mark.parametrize("nbsym", [1, 3, 4])
def test_other_nbsym_parity(nbsym):
    """Mirrors of 1, 3 and 4 points per side: the general mirror path (nbsym = 2 runs a
    constant-folded instance of it)."""
    from cuda_emd import EMDKernelConfig, synth
    x = np.concatenate([synth.chirp_mix(12, 1024, seed=4),
                        synth.seismic_traces(4, 2001, seed=5)[:, :1024]])
    o = [O.emd(s, max_imf=15, nbsym=nbsym) for s in x]
    _assert_parity(x, _gpu_padded(x, EMDKernelConfig(nbsym=nbsym)), o)


def test_overflow_small_max_extrema():
    from cuda_emd import EMDKernelConfig, synth
    x = synth.chirp_mix()[:16]
    cfg = EMDKernelConfig(max_extrema=24)
    g = _gpu_padded(x, cfg)
    o = _oracle(x, max_extrema=24)
    assert g["overflow"].sum() > 0
    _assert_parity(x, g, o)


def test_max_iteration_cap():
    from cuda_emd import EMDKernelConfig, synth
    x = synth.chirp_mix()[:16]
    cfg = EMDKernelConfig(max_iteration=2)
    g = _gpu_padded(x, cfg)
    o = _oracle(x, max_iteration=2)
    assert g["overflow"].all()
    _assert_parity(x, g, o)


def test_envelope_only_parity():
    import torch
    from cuda_emd import EMDKernelConfig, synth
    x = synth.chirp_mix()[:32]
    L = EMDKernelConfig().build()
    out = [t.cpu().numpy() for t in L.run_envelope_only(torch.from_numpy(x).cuda())]
    max_env, min_env, n_ext, degen, mkp, mkv, nkp, nkv, n_mir = out
    for b in range(len(x)):
        e = O.envelopes(x[b])
        assert degen[b] == e["degenerate"]
        assert tuple(n_ext[b]) == (e["n_max"], e["n_min"])
        km, kn = e["max_knots"], e["min_knots"]
        assert tuple(n_mir[b]) == (km.shape[1], kn.shape[1])
        assert np.array_equal(mkp[b, :km.shape[1]], km[0])
        assert np.array_equal(mkv[b, :km.shape[1]], km[1])
        assert np.array_equal(nkp[b, :kn.shape[1]], kn[0])
        assert np.array_equal(nkv[b, :kn.shape[1]], kn[1])
        scale = np.max(np.abs(x[b]))
        assert np.max(np.abs(max_env[b] - e["max_env"])) <= 1e-12 * scale
        assert np.max(np.abs(min_env[b] - e["min_env"])) <= 1e-12 * scale


def _walk(seed):
    rng = np.random.default_rng(seed)
    N = int(rng.integers(8, 40))
    return np.cumsum(rng.standard_normal(N))


@pytest.mark.parametrize("seed", [27, 353])
def test_imf_cap_with_trend_drop(seed):
    """Regression: the IMF cap reached while the last IMF has <= 2 extrema (PyEMD drops
    that row after leaving its loop).  Seeds found by scanning random walks with the
    oracle for exactly this path."""
    from cuda_emd import EMDKernelConfig
    x = _walk(seed)[None, :]
    o = _oracle(x, max_imfs=3)
    assert o[0].imf_cap_hit and 0 <= o[0].debug_extno[1] <= 2
    _assert_parity(x, _gpu_padded(x, EMDKernelConfig(max_imfs=3)), o, max_imfs=3)


@pytest.mark.parametrize("threads", [256, 1024])
def test_long_trace_parity(threads):
    """Traces of 6000-12000 samples (up to four 64-sample blocks per lane in the block
    scans, more than 64 block edges) through the persistent CSR path, with 256 and with
    1024 lanes per trace."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    rng = np.random.default_rng(21)
    sig = [synth.chirp_mix(1, 6000, seed=12)[0],
           synth.chirp_mix(1, 9000, seed=13)[0] + 0.02 * rng.standard_normal(9000),
           synth.seismic_traces(1, 12000, seed=7)[0]]
    launcher = EMDKernelConfig(max_extrema=4096).build()
    offsets = np.zeros(len(sig) + 1, np.int64)
    offsets[1:] = np.cumsum([len(s) for s in sig])
    out = launcher.run_ragged(torch.from_numpy(np.concatenate(sig)).cuda(),
                              torch.from_numpy(offsets), threads=threads, tile_slots=786,
                              persistent=True)
    torch.cuda.synchronize()
    flat = out[0].cpu().numpy()
    gpu = dict(imfs=[flat[offsets[b] * 16: offsets[b] * 16 + 16 * len(s)].reshape(16, len(s))
                     for b, s in enumerate(sig)],
               n_imfs=out[2].cpu().numpy(), overflow=out[3].cpu().numpy(),
               debug_n=out[4].cpu().numpy(), debug_extno=out[5].cpu().numpy(),
               debug_nzm=out[6].cpu().numpy(), stats=out[7].cpu().numpy())
    _assert_parity(sig, gpu, _oracle(sig, max_extrema=4096))
