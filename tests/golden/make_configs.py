"""Writes tests/golden/configs_sample.npz: stratified samples of three benchmark workloads,
members of the very batches bench.py times, with the oracle's records for them.

* ragged: synth.ragged_mix(65536, seed=2) -- 2 traces from each (content kind x length
  bucket) stratum: seismic / AR(1) 0.95 / gated sinusoid x N in (256-512], ..., (4096-8192];
  max_extrema 1024 (EMDKernelConfig.for_device on gfx950).
* heldout: synth.heldout_mix(262144, seed=4), the cost-model workload's held-out set --
  12 seismic traces (N=2001) and 5 white-noise traces (N=8192); max_extrema 1024.
* stress: synth.stress_mix(262144, 4096, seed=5) -- 6 chirp-mix, 4 AR(1) 0.99 and 3 FM
  chirp traces (N=4096); max_extrema 2048.

Stored per trace: workload, content kind, index in its batch, max_extrema, the samples
(CSR), and the oracle's (oracle/pyemd_restated.py, PyEMD EMD() defaults, IMF cap 15)
n_imfs, overflow, debug_n / debug_extno / debug_nzm, total sifts, peak extrema per side,
knot sum and near ties.  The IMF rows are not stored (about 10 rows per trace); the tests
recompute them with the oracle.  The reference ships no vectors for these workloads.

Run from the repository root (needs ~20 GB of memory and a few minutes):
    python tests/golden/make_configs.py
"""
import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "emd-cuda_amd"))

from cuda_emd import synth  # noqa: E402
from oracle import pyemd_restated as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "configs_sample.npz")
WORKLOADS = ("ragged", "heldout", "stress")
TIE_KINDS = ("sample_sign", "difference_sign", "total_power_1e-10", "svar", "std",
             "energy_ratio", "range", "total_power")
BUCKETS = (512, 1024, 2048, 4096, 8192)


def _pick(rng, idx, k):
    idx = np.asarray(idx)
    return np.sort(rng.choice(idx, min(k, len(idx)), replace=False)) if len(idx) else idx


def sample_ragged(rng):
    values, offsets, kinds = synth.ragged_mix(65536, seed=2)
    lens = np.diff(offsets)
    bucket = np.searchsorted(np.asarray(BUCKETS), lens, side="left")
    pick = []
    for kd in (0, 1, 2):
        for b in range(len(BUCKETS)):
            pick += list(_pick(rng, np.nonzero((kinds == kd) & (bucket == b))[0], 2))
    pick = np.array(sorted(pick))
    return [values[offsets[s]:offsets[s + 1]].copy() for s in pick], kinds[pick], pick


def sample_heldout(rng):
    values, offsets, kinds = synth.heldout_mix(262144, seed=4)
    pick = np.sort(np.r_[_pick(rng, np.nonzero(kinds == 0)[0], 12),
                         _pick(rng, np.nonzero(kinds == 1)[0], 5)])
    return [values[offsets[s]:offsets[s + 1]].copy() for s in pick], kinds[pick], pick


def sample_stress(rng):
    x, kinds = synth.stress_mix(262144, 4096, seed=5)
    pick = np.sort(np.r_[_pick(rng, np.nonzero(kinds == 0)[0], 6),
                         _pick(rng, np.nonzero(kinds == 1)[0], 4),
                         _pick(rng, np.nonzero(kinds == 2)[0], 3)])
    return [x[s].copy() for s in pick], kinds[pick], pick


def _oracle(args):
    x, me = args
    return O.emd(x, max_imf=15, max_extrema=int(me), debug_slots=16)


def main():
    rng = np.random.default_rng(2026)
    traces, wl, kind, index, me = [], [], [], [], []
    for w, (fn, max_ext) in enumerate(((sample_ragged, 1024), (sample_heldout, 1024),
                                       (sample_stress, 2048))):
        tr, kd, ix = fn(rng)
        print(WORKLOADS[w], len(tr), "traces", flush=True)
        traces += tr
        wl += [w] * len(tr)
        kind += list(kd)
        index += list(ix)
        me += [max_ext] * len(tr)
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        res = pool.map(_oracle, list(zip(traces, me)), chunksize=1)
    offsets = np.zeros(len(traces) + 1, np.int64)
    offsets[1:] = np.cumsum([len(t) for t in traces])
    ties = [(t_i, t.imf, t.sift, TIE_KINDS.index(t.what)) for t_i, r in enumerate(res)
            for t in r.near_ties]
    ties = np.array(ties, np.int64).reshape(-1, 4)
    np.savez_compressed(
        OUT, workload=np.array(wl, np.int8), kind=np.array(kind, np.int8),
        index=np.array(index, np.int64), max_extrema=np.array(me, np.int32),
        values=np.concatenate(traces), offsets=offsets,
        n_imfs=np.array([r.n_imfs for r in res], np.int32),
        overflow=np.array([r.overflow for r in res], np.int32),
        debug_n=np.array([r.debug_n for r in res], np.int32),
        debug_extno=np.array([r.debug_extno for r in res], np.int32),
        debug_nzm=np.array([r.debug_nzm for r in res], np.int32),
        counters=np.array([[r.total_sifts, r.peak_extrema, r.knot_sum] for r in res], np.int64),
        ties=ties)
    print("wrote", OUT, len(traces), "traces,", int(offsets[-1]), "samples,", len(ties), "ties")


if __name__ == "__main__":
    main()
