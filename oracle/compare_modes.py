"""How often the reference kernel's own rules and PyEMD's give different decompositions.

Runs the oracle (oracle/pyemd_restated.py) in its two modes -- PyEMD EMD() defaults, and
mode='reference' with the reference kernel's extrema/zero-crossing scan, 3-knot parabola
and running residue (reference kernels.py:115-349, 711-718, 994-1075) -- on samples of the
five benchmark workloads, and prints, per workload, how many traces differ in n_imfs, in
the per-IMF sift counts (debug_n), in any integer record, and the largest IMF difference
(x max|x|) among traces whose records agree.  Test infrastructure: it only compares the
oracle with itself.

Run from the repository root (a few minutes; about 30 GB of memory for the stress batch):
    python oracle/compare_modes.py [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "emd-cuda_amd"))

from cuda_emd import synth  # noqa: E402
from oracle import pyemd_restated as O  # noqa: E402


def samples():
    """(workload, description, traces, max_extrema) with seeded random members."""
    rng = np.random.default_rng(7)
    out = [("chirp mix", "256 x N=1024, seed 0 (all)", list(synth.chirp_mix(256, 1024, seed=0)),
            1024)]
    x = synth.seismic_traces(16384, 2001, seed=1)
    out.append(("seismic", "first 512 of 16384 x N=2001, seed 1", list(x[:512]), 1024))
    del x
    v, off, _ = synth.ragged_mix(65536, seed=2)
    pick = np.sort(rng.choice(65536, 600, replace=False))
    out.append(("ragged mix", "600 random of 65536 (N 256..8192), seed 2",
                [v[off[s]:off[s + 1]].copy() for s in pick], 1024))
    del v, off
    v, off, _ = synth.heldout_mix(262144, seed=4)
    pick = np.sort(rng.choice(262144, 600, replace=False))
    out.append(("cost-model held-out", "600 random of 262144 (5% white N=8192), seed 4",
                [v[off[s]:off[s + 1]].copy() for s in pick], 1024))
    del v, off
    x, _ = synth.stress_mix(262144, 4096, seed=5)
    pick = np.sort(rng.choice(262144, 400, replace=False))
    out.append(("stress", "400 random of 262144 x N=4096, seed 5", [x[s].copy() for s in pick],
                2048))
    del x
    return out


def _both(args):
    x, me = args
    a = O.emd(x, max_imf=15, max_extrema=me)
    b = O.emd(x, max_imf=15, max_extrema=me, mode="reference")
    rec_a = (a.n_imfs, a.overflow, tuple(a.debug_n), tuple(a.debug_extno), tuple(a.debug_nzm))
    rec_b = (b.n_imfs, b.overflow, tuple(b.debug_n), tuple(b.debug_extno), tuple(b.debug_nzm))
    err = 0.0
    if rec_a == rec_b and a.n_imfs:
        err = float(np.max(np.abs(a.imfs - b.imfs)) / max(np.max(np.abs(x)), 1e-300))
    return (a.n_imfs != b.n_imfs, tuple(a.debug_n) != tuple(b.debug_n), rec_a != rec_b, err)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--json", default=None)
    args = ap.parse_args()
    rows = []
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for name, desc, traces, me in samples():
            r = pool.map(_both, [(t, me) for t in traces], chunksize=4)
            rows.append({"workload": name, "sample": desc, "traces": len(traces),
                         "n_imfs_differ": sum(q[0] for q in r),
                         "debug_n_differ": sum(q[1] for q in r),
                         "any_record_differs": sum(q[2] for q in r),
                         "max_imf_diff_when_records_agree": max(q[3] for q in r)})
            print(json.dumps(rows[-1]), flush=True)
    print("\n| workload | sample | traces | n_imfs differ | debug_n differ | any record differs "
          "| max IMF diff (x max|x|) where records agree |")
    print("|---|---|---|---|---|---|---|")
    for q in rows:
        print(f"| {q['workload']} | {q['sample']} | {q['traces']} | {q['n_imfs_differ']} | "
              f"{q['debug_n_differ']} | {q['any_record_differs']} | "
              f"{q['max_imf_diff_when_records_agree']:.1e} |")
    if args.json:
        with open(args.json, "w") as f:
            json.dump(rows, f, indent=1)


if __name__ == "__main__":
    main()
