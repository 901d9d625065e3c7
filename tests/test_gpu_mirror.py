"""The envelope-only kernel (reference envelope_batch_kernel, kernels.py:1100-1166) on the
hand-derived mirror cases, on degenerate traces and with the max_extrema clamp."""
import numpy as np
import pytest

from oracle import pyemd_restated as O
from test_oracle_pinning import MIRROR_CASES

pytestmark = pytest.mark.gpu


def _env(x, cfg=None):
    import torch
    from cuda_emd import EMDKernelConfig
    L = (cfg or EMDKernelConfig()).build()
    out = L.run_envelope_only(torch.from_numpy(np.ascontiguousarray(np.atleast_2d(x))).cuda())
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in out]


@pytest.mark.parametrize("case", MIRROR_CASES, ids=[c[0] for c in MIRROR_CASES])
def test_mirror_branches_gpu(case):
    """Knot positions and values of every mirror branch equal the hand-derived lists."""
    name, S, (mp, mv), (np_, nv) = case
    max_env, min_env, n_ext, degen, mkp, mkv, nkp, nkv, n_mir = _env(np.asarray(S, np.float64))
    assert degen[0] == 0, name
    assert tuple(n_mir[0]) == (len(mp), len(np_)), name
    assert np.array_equal(mkp[0, :len(mp)], mp) and np.array_equal(mkv[0, :len(mv)], mv), name
    assert np.array_equal(nkp[0, :len(np_)], np_) and np.array_equal(nkv[0, :len(nv)], nv), name
    e = O.envelopes(np.asarray(S, np.float64))
    scale = np.max(np.abs(S))
    assert np.max(np.abs(max_env[0] - e["max_env"])) <= 1e-12 * scale, name
    assert np.max(np.abs(min_env[0] - e["min_env"])) <= 1e-12 * scale, name


def _degenerate_signals():
    n = np.arange(40, dtype=np.float64)
    return {
        "const": np.full(40, 2.5),
        "ramp": 0.1 * n - 1.0,
        "spike_on_ramp": np.where(n == 17, 1.0, 0.0) + 1e-3 * n,  # one max, one min
        "one_max": np.sin(2 * np.pi * n / 60.0),                 # one max, no min
    }


@pytest.mark.parametrize("name", list(_degenerate_signals()))
def test_envelope_degenerate(name):
    """extNo <= 2 or a side without extrema: degenerate = 1, the raw counts are reported and
    every other output keeps the caller's zeros (reference kernels.py:1133-1140)."""
    x = _degenerate_signals()[name]
    e = O.envelopes(x)
    assert e["degenerate"] == 1
    max_env, min_env, n_ext, degen, mkp, mkv, nkp, nkv, n_mir = _env(x)
    assert degen[0] == 1
    assert tuple(n_ext[0]) == (e["n_max"], e["n_min"])
    for a in (max_env, min_env, mkp, mkv, nkp, nkv, n_mir):
        assert not np.any(a), name


@pytest.mark.parametrize("N", [1, 2])
def test_envelope_too_short(N):
    x = np.arange(3 * N, dtype=np.float64).reshape(3, N) * np.array([[1.0], [-1.0], [0.5]])
    max_env, min_env, n_ext, degen, mkp, mkv, nkp, nkv, n_mir = _env(x)
    assert np.all(degen == 1) and not np.any(n_ext)
    assert not np.any(max_env) and not np.any(min_env) and not np.any(n_mir)


def test_envelope_max_extrema_clamp():
    """With max_extrema = 6 only the first six maxima and minima become knots, as in the
    reference's capacity rule; knots and envelopes match the oracle's clamped lists."""
    from cuda_emd import EMDKernelConfig, synth
    x = synth.chirp_mix(6, 300, seed=41)
    out = _env(x, EMDKernelConfig(max_extrema=6))
    max_env, min_env, n_ext, degen, mkp, mkv, nkp, nkv, n_mir = out
    for b in range(len(x)):
        e = O.envelopes(x[b], max_extrema=6)
        assert degen[b] == e["degenerate"] == 0
        assert tuple(n_ext[b]) == (e["n_max"], e["n_min"]) == (6, 6)
        km, kn = e["max_knots"], e["min_knots"]
        assert tuple(n_mir[b]) == (km.shape[1], kn.shape[1])
        assert np.array_equal(mkp[b, :km.shape[1]], km[0]) and np.array_equal(mkv[b, :km.shape[1]], km[1])
        assert np.array_equal(nkp[b, :kn.shape[1]], kn[0]) and np.array_equal(nkv[b, :kn.shape[1]], kn[1])
        assert not np.any(mkp[b, km.shape[1]:]) and not np.any(nkp[b, kn.shape[1]:])
        scale = np.max(np.abs(x[b]))
        # the envelopes are evaluated on the samples between the first and last knot
        assert np.max(np.abs(max_env[b] - e["max_env"])) <= 1e-9 * scale
        assert np.max(np.abs(min_env[b] - e["min_env"])) <= 1e-9 * scale
