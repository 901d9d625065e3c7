"""Replays, in host emulation, the persistent-kernel calls that hung on the GPU with two
earlier sift kernels, and the same calls on the current kernel (tests/native/sift_emu).

* 3202379 (samples in registers, LDS knot tile, deferral + retry launch):
  test_persistent_ragged_matches_plain's three calls on synth.ragged_mix(B=48, seed=11,
  100..3000 samples): persistent 256 lanes; plain 128 lanes; persistent 128 lanes with a
  300-slot tile and a random order (traces deferred to the retry launch).
* 97b620f (LDS-resident IMF, deferral + retry launch over a global tile):
  test_persistent_ragged_seismic_chirp's third call (24 seismic / chirp traces, 64 lanes, a
  200-slot tile, a random order).

Each call runs with as many concurrent workgroups as traces (the grid the GPU launch had),
and is held to the oracle parity bar.  Build first: make -C tests/native/history."""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", "..", ".."))
sys.path[:0] = [os.path.join(ROOT, "tests"), os.path.join(ROOT, "emd-cuda_amd"), ROOT]

from test_sift_emu import _assert_parity, _ragged_mix_48, run_emu  # noqa: E402


def seismic_chirp_24():
    from cuda_emd import synth
    rng = np.random.default_rng(21)
    seis = synth.seismic_traces(12, 2001, seed=21)
    chirp = synth.chirp_mix(12, 1500, seed=22)
    traces = []
    for k in range(24):
        base = seis[k // 2] if k % 2 == 0 else chirp[k // 2]
        L = int(rng.integers(200, base.size + 1))
        traces.append(np.ascontiguousarray(base[:L]))
    return traces, rng.permutation(24).astype(np.int32)


def main():
    cur = os.path.join(ROOT, "tests", "native", "sift_emu")
    ragged = _ragged_mix_48()
    order48 = np.random.default_rng(0).permutation(48).astype(np.int32)
    sc, order24 = seismic_chirp_24()
    calls = [
        ("persistent 256", ragged, dict(wg=256, persistent=True)),
        ("plain 128", ragged, dict(wg=128, persistent=False)),
        ("persistent 128, tile 300, order", ragged, dict(wg=128, tile=300, persistent=True,
                                                         order=order48)),
        ("persistent 64, tile 200, order", sc, dict(wg=64, tile=200, persistent=True,
                                                    order=order24)),
    ]
    runs = [("3202379", os.path.join(HERE, "build", "emu_3202379"), calls[:3]),
            ("97b620f", os.path.join(HERE, "build", "emu_97b620f"), calls[3:]),
            ("current", cur, calls)]
    for name, exe, cs in runs:
        for label, traces, kw in cs:
            kw = dict(kw)
            wg = kw.pop("wg")
            t0 = time.time()
            out = run_emu(exe, traces, wg, resident=len(traces), timeout=3000, **kw)
            _assert_parity(traces, out, f"{name}:{label}")
            print(f"{name:8s} {label:34s} done in {time.time() - t0:5.1f} s, "
                  f"{int(np.sum(out['stats'][:, 3] & 2 > 0))} traces retried or on the global "
                  f"tile, oracle parity", flush=True)


if __name__ == "__main__":
    main()
