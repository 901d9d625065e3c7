"""Shared parity checks: GPU outputs against the CPU oracle (oracle/pyemd_restated.py).

Bar: n_imfs, overflow, the per-IMF records (debug_n sift counts, debug_extno,
debug_nzm) and the kernel's own counters (total sifts, peak extrema per side, knot sum)
identical; IMF samples within 1e-9 * max|x| (max-abs).

A mismatch of the integer records is excused only when the oracle logged a near tie (a
stop statistic within 1e-12 of its threshold, or a sign test within 1e-12 * max|imf| of
zero) where the two decompositions part ways: in the first IMF slot whose records differ,
at a sift no later than the smaller of the two sift counts there, or the end-condition
test of the IMF just before it.  IMF value errors are never excused when every integer
record matches.  Every excused divergence is appended, with the ties that excuse it, to
a JSON-lines log (``EMD_PARITY_LOG``, default ``gpurun_out/parity_divergences.jsonl``).
"""
from __future__ import annotations

import json
import os

import numpy as np

IMF_TOL = 1e-9
# worst IMF error (x max|x|) of the last compare(), over all traces and over the traces
# whose decomposition did not overflow (an overflowed decomposition keeps clamped knot
# lists, whose far-reaching mirrored knots can make IMFs hundreds of times max|x|)
LAST_WORST = {"all": 0.0, "no_overflow": 0.0}
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LOG_PATH = os.environ.get("EMD_PARITY_LOG",
                          os.path.join(ROOT, "gpurun_out", "parity_divergences.jsonl"))
END_TIES = ("range", "total_power")


def first_divergence(g_rec, o_rec, g_n, o_n, max_imfs):
    """(k, s): the first IMF slot whose per-IMF records differ (or the shorter
    decomposition's length when they agree on every slot both reached) and the smaller
    of the two sift counts there (-1 when neither side reached the slot)."""
    diff = np.zeros(max_imfs, bool)
    for g, o in zip(g_rec, o_rec):
        diff |= np.asarray(g)[:max_imfs] != np.asarray(o)[:max_imfs]
    idx = np.nonzero(diff)[0]
    k = int(idx[0]) if idx.size else int(min(g_n, o_n))
    if k >= max_imfs:
        return k, -1
    cands = [int(v) for v in (g_rec[0][k], o_rec[0][k]) if int(v) >= 1]
    return k, (min(cands) if cands else -1)


def excusing_ties(ties, k, s):
    """The oracle's near ties that can explain a divergence first seen at (k, s)."""
    out = []
    for t in ties:
        at_k = t.imf == k and (s < 0 or t.sift <= s)
        end_before = t.imf == k - 1 and t.what in END_TIES
        if at_k or end_before:
            out.append(t)
    return out


def _log(entry):
    try:
        os.makedirs(os.path.dirname(LOG_PATH), exist_ok=True)
        with open(LOG_PATH, "a") as f:
            f.write(json.dumps(entry) + "\n")
    except OSError:
        pass


def compare(signals, gpu, ora, max_imfs=16, label=""):
    """signals: list of 1-D arrays; gpu: dict of numpy arrays (imfs: a list of [rows, N]
    arrays or the padded array; n_imfs, overflow, debug_n, debug_extno, debug_nzm; stats
    [B,4] optional); ora: oracle results (EMDResult list).  Returns (mismatches,
    worst_rel_err); a mismatch is (index, field, explained_by_near_tie)."""
    bad = []
    worst = worst_ok = 0.0
    stats = gpu.get("stats")
    for b, x in enumerate(signals):
        r = ora[b]
        g_rec = [np.asarray(gpu[k][b]) for k in ("debug_n", "debug_extno", "debug_nzm")]
        o_rec = [np.asarray(v)[:max_imfs] for v in (r.debug_n, r.debug_extno, r.debug_nzm)]
        fields = []
        if int(gpu["n_imfs"][b]) != r.n_imfs:
            fields.append("n_imfs")
        if int(gpu["overflow"][b]) != r.overflow:
            fields.append("overflow")
        for name, g, o in zip(("debug_n", "debug_extno", "debug_nzm"), g_rec, o_rec):
            if not np.array_equal(g, o):
                fields.append(name)
        if stats is not None:
            st = np.asarray(stats[b])
            for j, (name, ref) in enumerate((("sifts", r.total_sifts),
                                             ("peak_extrema", r.peak_extrema),
                                             ("knot_sum", r.knot_sum))):
                if int(st[j]) != int(ref):
                    fields.append(f"{name}({int(st[j])}!={int(ref)})")
        integer_ok = not fields
        if integer_ok:
            k = r.n_imfs
            g = np.asarray(gpu["imfs"][b])[:k]
            scale = float(np.max(np.abs(x))) if x.size else 0.0
            scale = scale if scale > 0 else 1.0
            err = float(np.max(np.abs(g - r.imfs))) / scale if k else 0.0
            worst = max(worst, err)
            if not r.overflow:
                worst_ok = max(worst_ok, err)
            if not (err <= IMF_TOL):
                fields.append(f"imfs(err={err:.3e})")
                for f in fields:
                    bad.append((b, f, False))
            continue
        k, s = first_divergence(g_rec, o_rec, int(gpu["n_imfs"][b]), r.n_imfs, max_imfs)
        ties = excusing_ties(r.near_ties, k, s)
        explained = bool(ties)
        for f in fields:
            bad.append((b, f, explained))
        _log({"label": label, "trace": b, "N": int(x.size), "fields": fields,
              "first_divergence": [k, s], "explained": explained,
              "gpu": {"n_imfs": int(gpu["n_imfs"][b]), "debug_n": g_rec[0].tolist()},
              "oracle": {"n_imfs": r.n_imfs, "debug_n": o_rec[0].tolist()},
              "ties": [[t.imf, t.sift, t.what, t.value, t.threshold] for t in ties[:8]]})
    LAST_WORST["all"], LAST_WORST["no_overflow"] = worst, worst_ok
    return bad, worst
