"""Golden fixtures (tests/golden/emd_small.npz, written by tests/golden/make_golden.py):
the oracle must still reproduce them exactly (CPU: a regression guard for the oracle, whose
own correctness is pinned by tests/test_oracle.py), and the HIP path must match them to
the parity bar (GPU) -- exactly where the oracle logged no near tie, and where it did, by
the parity rule that excuses a difference only at a tie where the two decompositions part
ways.  Every case must also add up to its trace."""
import os

import numpy as np
import pytest

from oracle import pyemd_restated as O
from parity import compare

GOLD = os.path.join(os.path.dirname(__file__), "golden", "emd_small.npz")


def _load():
    g = np.load(GOLD, allow_pickle=False)
    off = np.zeros(len(g["lengths"]) + 1, np.int64)
    off[1:] = np.cumsum(g["lengths"])
    return g, off


def _rows(g, k, off_rows):
    n, L = int(g["n_imfs"][k]), int(g["lengths"][k])
    return g["rows"][off_rows[k]:off_rows[k] + n * L].reshape(n, L)


def _row_offsets(g):
    ro = np.zeros(len(g["lengths"]) + 1, np.int64)
    ro[1:] = np.cumsum(g["n_imfs"].astype(np.int64) * g["lengths"])
    return ro


def test_oracle_reproduces_golden():
    g, off = _load()
    ro = _row_offsets(g)
    for k in range(len(g["lengths"])):
        x = g["values"][off[k]:off[k + 1]]
        r = O.emd(x, max_imf=15)
        assert r.n_imfs == g["n_imfs"][k] and r.overflow == g["overflow"][k], g["names"][k]
        assert np.array_equal(r.debug_n[:16], g["debug_n"][k])
        assert np.array_equal(r.debug_extno[:16], g["debug_extno"][k])
        assert np.array_equal(r.debug_nzm[:16], g["debug_nzm"][k])
        assert np.array_equal(r.imfs, _rows(g, k, ro))


@pytest.mark.gpu
def test_gpu_matches_golden():
    import torch
    from cuda_emd import EMDKernelConfig
    g, off = _load()
    ro = _row_offsets(g)
    launcher = EMDKernelConfig().build()
    out = launcher.run_ragged(torch.from_numpy(g["values"]).cuda(), torch.from_numpy(off),
                              threads=256, persistent=True)
    imfs, oo = out[0].cpu().numpy(), out[1].cpu().numpy()
    for k in range(len(g["lengths"])):
        name = str(g["names"][k])
        x = g["values"][off[k]:off[k + 1]]
        L = int(g["lengths"][k])
        scale = max(float(np.max(np.abs(x))), 1e-300) if L else 1.0
        n = int(out[2][k])
        rows = imfs[oo[k]:oo[k] + n * L].reshape(n, L)
        # every decomposition adds up to its trace (the residue row is dropped only when
        # max|residue| <= 1e-8)
        assert np.max(np.abs(rows.sum(axis=0) - x), initial=0.0) <= 1e-8 + 1e-9 * scale, name
        if g["near_ties"][k]:
            # the oracle logged near ties here: the parity rule decides whether a
            # difference is excused (only by a tie where the two decompositions part ways)
            gpu = dict(imfs=[rows], n_imfs=[n], overflow=[int(out[3][k])],
                       debug_n=[out[4][k].cpu().numpy()], debug_extno=[out[5][k].cpu().numpy()],
                       debug_nzm=[out[6][k].cpu().numpy()], stats=[out[7][k].cpu().numpy()])
            bad, _ = compare([x], gpu, [O.emd(x, max_imf=15)], label=f"golden:{name}")
            assert all(m[2] for m in bad), (name, bad)
            continue
        assert int(out[2][k]) == g["n_imfs"][k], name
        assert int(out[3][k]) == g["overflow"][k], name
        for j, key in ((4, "debug_n"), (5, "debug_extno"), (6, "debug_nzm")):
            assert np.array_equal(out[j][k].cpu().numpy(), g[key][k]), (name, key)
        n = int(g["n_imfs"][k])
        got = imfs[oo[k]:oo[k] + n * L].reshape(n, L)
        assert np.max(np.abs(got - _rows(g, k, ro)), initial=0.0) <= 1e-9 * scale, name


CONFIGS = os.path.join(os.path.dirname(__file__), "golden", "configs_sample.npz")
TIE_KINDS = ("sample_sign", "difference_sign", "total_power_1e-10", "svar", "std",
             "energy_ratio", "range", "total_power")


def test_oracle_reproduces_config_sample():
    """The oracle still gives the records frozen for the ragged-mix, held-out and stress
    samples (tests/golden/make_configs.py), near ties included."""
    g = np.load(CONFIGS, allow_pickle=False)
    off = g["offsets"]
    ties = g["ties"]
    for s in range(len(g["n_imfs"])):
        r = O.emd(g["values"][off[s]:off[s + 1]], max_imf=15,
                  max_extrema=int(g["max_extrema"][s]), debug_slots=16)
        assert r.n_imfs == g["n_imfs"][s] and r.overflow == g["overflow"][s], s
        assert np.array_equal(r.debug_n, g["debug_n"][s]), s
        assert np.array_equal(r.debug_extno, g["debug_extno"][s]), s
        assert np.array_equal(r.debug_nzm, g["debug_nzm"][s]), s
        assert [r.total_sifts, r.peak_extrema, r.knot_sum] == g["counters"][s].tolist(), s
        mine = sorted((t.imf, t.sift, TIE_KINDS.index(t.what)) for t in r.near_ties)
        frozen = sorted(map(tuple, ties[ties[:, 0] == s][:, 1:].tolist()))
        assert mine == frozen, s
