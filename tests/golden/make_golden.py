"""Writes the committed golden fixtures: small inputs and the oracle's outputs for them
(PyEMD EMD() defaults, restated in oracle/pyemd_restated.py).

The reference ships no golden vectors, fixtures or known-answer tests for this path, and
neither PyEMD nor the reference (CuPy/CUDA) can run in this image, so these fixtures are
the oracle's own outputs, frozen: a CPU test checks that the oracle still reproduces them
(any change to the oracle shows up as a diff here) and a GPU test checks the HIP path
against them without running the oracle.  ``near_ties`` counts the stop statistics and
sign tests the oracle found within 1e-12 of their threshold for each case; where it is
non-zero, rounding-level differences between two correct implementations may change the
counts (e.g. the flat steps of "stairs", whose mean envelope is zero up to rounding), so
the GPU test holds those cases to the reconstruction identity instead.

Run from the repository root:  python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "emd-cuda_amd"))

from cuda_emd import synth  # noqa: E402
from oracle import pyemd_restated as O  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "emd_small.npz")


def cases():
    """(name, signal) pairs: config-1 chirp mixes, seismic traces, edge signals."""
    out = [(f"chirp{k}", x) for k, x in enumerate(synth.chirp_mix(4, 400, seed=0))]
    out += [(f"seismic{k}", x) for k, x in enumerate(synth.seismic_traces(3, 500, seed=1))]
    n = np.arange(300, dtype=np.float64)
    out += [("tone", np.sin(2 * np.pi * n / 23.0)),
            ("ramp_tone", 0.01 * n + np.sin(2 * np.pi * n / 17.0)),
            ("stairs", np.floor(n / 7.0) % 3 - 1.0),
            ("gated", synth.gated_sinusoid(300, np.random.default_rng(4))),
            ("zeros", np.zeros(64)),
            ("short5", np.array([0.0, 1.0, -1.0, 2.0, 0.5]))]
    return out


def main():
    names, values, lengths, rows = [], [], [], []
    n_imfs, overflow, dn, de, dz, ties = [], [], [], [], [], []
    for name, x in cases():
        r = O.emd(np.asarray(x, dtype=np.float64), max_imf=15)
        names.append(name)
        values.append(np.asarray(x, dtype=np.float64))
        lengths.append(len(x))
        rows.append(r.imfs.reshape(-1))
        n_imfs.append(r.n_imfs)
        overflow.append(r.overflow)
        dn.append(r.debug_n[:16])
        de.append(r.debug_extno[:16])
        dz.append(r.debug_nzm[:16])
        ties.append(len(r.near_ties))
    np.savez_compressed(
        OUT, names=np.array(names), values=np.concatenate(values),
        lengths=np.array(lengths, np.int64), rows=np.concatenate(rows),
        n_imfs=np.array(n_imfs, np.int32), overflow=np.array(overflow, np.int32),
        debug_n=np.array(dn, np.int32), debug_extno=np.array(de, np.int32),
        debug_nzm=np.array(dz, np.int32), near_ties=np.array(ties, np.int32))
    print("wrote", OUT, len(names), "cases")


if __name__ == "__main__":
    main()
