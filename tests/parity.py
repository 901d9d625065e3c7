"""Shared parity checks: GPU outputs against the CPU oracle (oracle/pyemd_restated.py).

Bar (north_star): n_imfs, per-IMF sift counts (debug_n), extrema counts (debug_extno,
debug_nzm) and overflow identical; IMF samples within 1e-9 * max|x| (max-abs); any
divergence must coincide with a stop statistic the oracle logged within 1e-12 of its
threshold.
"""
from __future__ import annotations

import numpy as np

IMF_TOL = 1e-9


def compare(signals, gpu, ora, max_imfs=16, ragged=None):
    """signals: list of 1-D arrays; gpu: dict of numpy arrays (imfs list per trace or
    padded array, n_imfs, overflow, debug_n, debug_extno, debug_nzm); ora: oracle results
    (list of EMDResult).  Returns (mismatches, worst_rel_err); a mismatch is
    (index, field, explained_by_near_tie)."""
    bad = []
    worst = 0.0
    for b, x in enumerate(signals):
        r = ora[b]
        fields = []
        if int(gpu["n_imfs"][b]) != r.n_imfs:
            fields.append("n_imfs")
        if int(gpu["overflow"][b]) != r.overflow:
            fields.append("overflow")
        for name, ref in (("debug_n", r.debug_n), ("debug_extno", r.debug_extno),
                          ("debug_nzm", r.debug_nzm)):
            if not np.array_equal(np.asarray(gpu[name][b]), np.asarray(ref[:max_imfs])):
                fields.append(name)
        if not fields:
            k = r.n_imfs
            g = gpu["imfs"][b][:k]
            scale = float(np.max(np.abs(x))) if x.size else 0.0
            scale = scale if scale > 0 else 1.0
            err = float(np.max(np.abs(g - r.imfs))) / scale if k else 0.0
            worst = max(worst, err)
            if not (err <= IMF_TOL):
                fields.append(f"imfs(err={err:.3e})")
        if fields:
            # first IMF slot where the decompositions part ways
            first = min(r.n_imfs, int(gpu["n_imfs"][b]))
            for name, ref in (("debug_n", r.debug_n), ("debug_extno", r.debug_extno),
                              ("debug_nzm", r.debug_nzm)):
                diff = np.nonzero(np.asarray(gpu[name][b]) != np.asarray(ref[:max_imfs]))[0]
                if diff.size:
                    first = min(first, int(diff[0]))
            explained = any(t.imf <= first for t in r.near_ties)
            for f in fields:
                bad.append((b, f, explained))
    return bad, worst
