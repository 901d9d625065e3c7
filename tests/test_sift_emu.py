"""The sift kernels' own source (emd-cuda_amd/csrc/emd_sift_dev.h) run on CPU threads by
tests/native/sift_emu: every lane a thread, barriers and wave-level operations emulated
(tests/native/emu/hip/hip_runtime.h), every knot-tile and work-area index bounds-checked
(EMD_BOUNDS_CHECK), under AddressSanitizer and UndefinedBehaviorSanitizer, with a time
limit on every barrier that reports where each lane stands if one is never completed.

The default cases run plain and persistent launches at 64 and 128 lanes, a knot tile small
enough that every sift spills to the workgroup's global tile, a random queue order, and edge
signals, and hold the outputs to the same parity bar as the GPU tests.  The three calls that
hung on the GPU in earlier builds (tests/test_gpu_identity.py, skipped there) are replayed
call by call, with the same workgroup sizes, tiles, queue orders and resident grids."""
import fcntl
import os
import subprocess
import sys
import tempfile

import numpy as np
import pytest

from oracle import pyemd_restated as O
from parity import compare

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
EMU = os.path.join(NATIVE, "sift_emu")
MAGIC = 0x454D5531


@pytest.fixture(scope="module")
def emu():
    # make rebuilds when the kernel source is newer; the lock keeps parallel test workers
    # from building at the same time
    with open(os.path.join(NATIVE, ".sift_emu.lock"), "w") as lk:
        fcntl.flock(lk, fcntl.LOCK_EX)
        subprocess.run(["make", "-s", "-C", NATIVE, "sift_emu"], check=True)
    return EMU


def run_emu(exe, traces, wg, tile=0, resident=2, persistent=True, order=None,
            max_extrema=1024, max_imfs=16, max_iteration=1000, timeout=1800, times=None,
            nbsym=2):
    """Decompose `traces` (list of 1-D float64) in the emulator; returns the parity dict.
    times: one normalised time axis per trace (the time-axis kernels), or None."""
    lengths = np.array([len(t) for t in traces], np.int64)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    values = np.concatenate(traces).astype(np.float64) if len(traces) else np.zeros(0)
    B, K = len(traces), max_imfs
    with tempfile.TemporaryDirectory() as d:
        fi, fo = os.path.join(d, "in"), os.path.join(d, "out")
        with open(fi, "wb") as f:
            np.array([MAGIC, B, wg, tile, resident, int(persistent), max_extrema, nbsym,
                      max_iteration, K, order is not None, times is not None],
                     np.int32).tofile(f)
            np.array([values.size], np.int64).tofile(f)
            offsets.tofile(f)
            values.tofile(f)
            if order is not None:
                np.asarray(order, np.int32).tofile(f)
            if times is not None:
                np.concatenate([np.asarray(t, np.float64) for t in times]).tofile(f)
        env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:handle_abort=1",
                   UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
        r = subprocess.run([exe, fi, fo], capture_output=True, text=True, timeout=timeout,
                           env=env)
        assert r.returncode == 0, r.stderr[-4000:]
        raw = np.fromfile(fo, np.uint8)
    nint = 2 * B + 3 * B * K + 4 * B
    ints = raw[:4 * nint].view(np.int32)
    imfs = raw[4 * nint:].view(np.float64)
    out = {"n_imfs": ints[:B], "overflow": ints[B:2 * B]}
    p = 2 * B
    for name in ("debug_n", "debug_extno", "debug_nzm"):
        out[name] = ints[p:p + B * K].reshape(B, K)
        p += B * K
    out["stats"] = ints[p:p + 4 * B].reshape(B, 4)
    out["imfs"] = [imfs[offsets[s] * K: offsets[s] * K + out["n_imfs"][s] * lengths[s]]
                   .reshape(out["n_imfs"][s], lengths[s]) for s in range(B)]
    return out


def _assert_parity(traces, out, label, max_extrema=1024, max_iteration=1000):
    ora = [O.emd(x, max_imf=15, max_extrema=max_extrema, max_iteration=max_iteration)
           for x in traces]
    bad, _ = compare(list(traces), out, ora, 16, label=label)
    unexplained = [m for m in bad if not m[2]]
    assert not unexplained, unexplained


def _same(a, b):
    for k in ("n_imfs", "overflow", "debug_n", "debug_extno", "debug_nzm"):
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["stats"][:, :3], b["stats"][:, :3])
    for x, y in zip(a["imfs"], b["imfs"]):
        assert np.array_equal(x, y)


def test_plain_and_persistent_match_oracle(emu):
    from cuda_emd import synth
    traces = list(synth.chirp_mix(3, 400, seed=5))
    plain = run_emu(emu, traces, 64, persistent=False, resident=3)
    pers = run_emu(emu, traces, 128, persistent=True, resident=2)
    _assert_parity(traces, plain, "emu:plain64")
    _same(plain, pers)


def test_global_tile_and_random_order(emu):
    """A 40-slot LDS tile sends every sift's knots to the workgroup's global tile; ragged
    lengths and a random order share one queue."""
    from cuda_emd import synth
    rng = np.random.default_rng(3)
    base = synth.chirp_mix(4, 520, seed=9)
    traces = [np.ascontiguousarray(base[k, :int(rng.integers(150, 521))]) for k in range(4)]
    order = rng.permutation(4).astype(np.int32)
    out = run_emu(emu, traces, 64, tile=40, resident=3, persistent=True, order=order)
    assert np.all(out["stats"][:, 3] & 2)  # every trace used the global tile
    _assert_parity(traces, out, "emu:global-tile")
    ref = run_emu(emu, traces, 64, persistent=False, resident=4)
    _same(out, ref)


def test_edge_signals(emu):
    """Empty, one- to five-sample, constant, zero, spike and square traces: no access
    outside any array, every barrier completed, outputs at the parity bar."""
    n = np.arange(130, dtype=np.float64)
    traces = [np.zeros(0), np.array([1.0]), np.array([1.0, -1.0]), np.array([0.0, 2.0, 0.0]),
              np.array([3.0, 1.0, 4.0, 1.0, 5.0]), np.full(70, 2.5), np.zeros(65),
              np.where(n % 17 == 0, 1.0, 0.0), np.sign(np.sin(2 * np.pi * n / 23.0))]
    out = run_emu(emu, traces, 64, resident=3, persistent=True)
    _assert_parity(traces, out, "emu:edge")


# ---- replays of the calls that hung on the GPU with earlier kernels ---------------------

def _ragged_mix_48():
    from cuda_emd import synth
    values, offsets, _ = synth.ragged_mix(B=48, seed=11, n_lo=100, n_hi=3000)
    return [values[offsets[s]:offsets[s + 1]].copy() for s in range(48)]


@pytest.mark.parametrize("call", ["persistent256", "plain128", "persistent128-tile300-order"])
def test_replay_persistent_ragged_matches_plain(emu, call):
    """tests/test_gpu_identity.py::test_persistent_ragged_matches_plain, call by call."""
    traces = _ragged_mix_48()
    if call == "persistent256":
        out = run_emu(emu, traces, 256, resident=4, persistent=True)
    elif call == "plain128":
        out = run_emu(emu, traces, 128, resident=4, persistent=False)
    else:
        order = np.random.default_rng(0).permutation(48).astype(np.int32)
        out = run_emu(emu, traces, 128, tile=300, resident=4, persistent=True, order=order)
    _assert_parity(traces, out, f"emu:replay-{call}")


def test_replay_persistent_ragged_seismic_chirp_small_tile(emu):
    """tests/test_gpu_identity.py::test_persistent_ragged_seismic_chirp, its third call:
    64-lane workgroups, a 200-slot tile, a random order."""
    from cuda_emd import synth
    rng = np.random.default_rng(21)
    seis = synth.seismic_traces(12, 2001, seed=21)
    chirp = synth.chirp_mix(12, 1500, seed=22)
    traces = []
    for k in range(24):
        base = seis[k // 2] if k % 2 == 0 else chirp[k // 2]
        L = int(rng.integers(200, base.size + 1))
        traces.append(np.ascontiguousarray(base[:L]))
    order = rng.permutation(24).astype(np.int32)
    out = run_emu(emu, traces, 64, tile=200, resident=6, persistent=True, order=order)
    _assert_parity(traces, out, "emu:replay-seismic-chirp")


def test_long_spline_systems(emu):
    """White-noise traces of 1600 samples: each side's spline system has hundreds of rows,
    so every lane's chunk has many interior rows whose factors the sweep parks in the
    chunk's knot slots."""
    rng = np.random.default_rng(7)
    traces = [rng.standard_normal(1600) for _ in range(2)]
    out = run_emu(emu, traces, 128, resident=2, persistent=True)
    _assert_parity(traces, out, "emu:long-systems")
    ref = run_emu(emu, traces, 64, resident=2, persistent=False)
    _same(out, ref)


def test_time_axis(emu):
    """The time-axis kernels (the reference's T: knots at T[idx], mirrored knots at
    2 T[sym] - T[idx], envelopes evaluated at T[i]) on jittered sampling, against the
    oracle given the same T: persistent at 64 lanes, and plain at 128 lanes with a 40-slot
    LDS tile so the knots go to the global tile of 32-byte slots."""
    from cuda_emd import synth
    rng = np.random.default_rng(11)
    traces = [np.ascontiguousarray(x) for x in synth.chirp_mix(3, 300, seed=13)]
    T = [np.cumsum(rng.uniform(0.6, 1.4, len(x))) for x in traces]
    ora = [O.emd(x, max_imf=15, max_extrema=1024, T=t) for x, t in zip(traces, T)]
    times = [O.normalize_time(t) for t in T]
    out = run_emu(emu, traces, 64, resident=2, persistent=True, times=times)
    bad, _ = compare(list(traces), out, ora, 16, label="emu:time-axis")
    assert not bad
    small = run_emu(emu, traces, 128, tile=40, resident=3, persistent=False, times=times)
    bad, _ = compare(list(traces), small, ora, 16, label="emu:time-axis-global-tile")
    assert not bad
    _same(out, small)


def test_time_axis_edge_signals(emu):
    """Plateaus, exact zeros and short traces on jittered axes in the emulation (the
    serial plateau rules, the small spline kinds and degenerate traces with time
    positions), against the oracle."""
    rng = np.random.default_rng(12)
    n = np.arange(257, dtype=np.float64)
    traces = [np.floor(np.sin(2 * np.pi * n / 50.0) * 4.0),
              np.where((n // 40) % 2 == 0, np.sin(2 * np.pi * 0.05 * n), 0.0),
              np.sign(np.sin(2 * np.pi * n / 30.0)),
              rng.standard_normal(5), rng.standard_normal(2), np.zeros(0)]
    T = [np.cumsum(rng.uniform(0.6, 1.4, len(x))) for x in traces]
    ora = [O.emd(x, max_imf=15, max_extrema=1024, T=t) for x, t in zip(traces, T)]
    times = [O.normalize_time(t) for t in T]
    out = run_emu(emu, traces, 64, resident=2, persistent=True, times=times)
    bad, _ = compare(list(traces), out, ora, 16, label="emu:time-axis-edge")
    assert not [m for m in bad if not m[2]], bad


@pytest.mark.parametrize("nbsym", [1, 3, 4])
def test_other_nbsym(emu, nbsym):
    """Mirrors of 1, 3 and 4 points per side (the general mirror path; nbsym = 2 has its
    own constant-folded instance), against the oracle."""
    from cuda_emd import synth
    traces = [np.ascontiguousarray(x) for x in synth.chirp_mix(3, 300, seed=17)]
    out = run_emu(emu, traces, 64, resident=2, persistent=True, nbsym=nbsym)
    ora = [O.emd(x, max_imf=15, max_extrema=1024, nbsym=nbsym) for x in traces]
    bad, _ = compare(list(traces), out, ora, 16, label=f"emu:nbsym{nbsym}")
    assert not [m for m in bad if not m[2]], bad
