"""Oracle parity on the ragged-mix, cost-model held-out and stress workloads: stratified
members of the benchmarked batches (tests/golden/configs_sample.npz, written by
tests/golden/make_configs.py) through the production launch shapes -- PackedEMD's length
buckets (persistent 256- and 1024-lane instances captured into graphs, one stream per
bucket) and StreamingEMD's chunk graphs (N=4096, max_extrema 2048, compacted rows).

The bar is the parity rule of tests/parity.py: n_imfs, overflow, the per-IMF records and
the kernel counters identical, IMF samples within 1e-9 * max|x|, a difference excused
only by an oracle near tie where the decompositions part ways."""
import os

import numpy as np
import pytest

from oracle import pyemd_restated as O
from parity import LAST_WORST, compare

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden", "configs_sample.npz")
RAGGED, HELDOUT, STRESS = 0, 1, 2


def _load(workloads):
    g = np.load(GOLD, allow_pickle=False)
    sel = np.nonzero(np.isin(g["workload"], workloads))[0]
    off = g["offsets"]
    traces = [g["values"][off[s]:off[s + 1]] for s in sel]
    return g, sel, traces


_ORACLE = {}  # fixture index -> oracle result (shared by the parametrised tests)


def _oracle(g, sel, traces):
    """Oracle results for the selected traces; their integer records must still be the
    ones frozen in the fixture."""
    res = []
    for s, x in zip(sel, traces):
        if s not in _ORACLE:
            _ORACLE[s] = O.emd(x, max_imf=15, max_extrema=int(g["max_extrema"][s]),
                               debug_slots=16)
        r = _ORACLE[s]
        assert r.n_imfs == g["n_imfs"][s] and r.overflow == g["overflow"][s], s
        assert np.array_equal(r.debug_n, g["debug_n"][s]), s
        assert [r.total_sifts, r.peak_extrema, r.knot_sum] == g["counters"][s].tolist(), s
        res.append(r)
    return res


def _assert_parity(label, traces, gpu, ora):
    bad, worst = compare(traces, gpu, ora, label=label)
    unexplained = [m for m in bad if not m[2]]
    print(f"{label}: {len(traces)} traces, {len(bad)} mismatches ({len(unexplained)} "
          f"unexplained), worst IMF err {worst:.3e} x max|x| "
          f"({LAST_WORST['no_overflow']:.3e} without overflow)")
    for m in bad:
        print("  mismatch", m)
    assert not unexplained, unexplained


class _FixedTuner:
    """Launch shapes per length bucket as the ragged benchmark's autotuner picks them:
    256 lanes up to 2048 samples, 1024 lanes (one workgroup of 16 waves per CU) above."""

    def __init__(self, small_tile=0):
        self.small_tile = small_tile

    def choose(self, launcher, values, offsets_host, idx, bucket, tile_hint=0):
        from cuda_emd.packed import LaunchChoice
        threads = 256 if bucket <= 2 else 1024
        return LaunchChoice(threads, self.small_tile, 0)


def _ragged_records(out, lengths):
    imfs, oo = out[0].cpu().numpy(), out[1].cpu().numpy()
    n = out[2].cpu().numpy()
    rows = [imfs[oo[s]:oo[s] + n[s] * lengths[s]].reshape(n[s], lengths[s])
            for s in range(len(lengths))]
    return dict(imfs=rows, n_imfs=n, overflow=out[3].cpu().numpy(),
                debug_n=out[4].cpu().numpy(), debug_extno=out[5].cpu().numpy(),
                debug_nzm=out[6].cpu().numpy(), stats=out[7].cpu().numpy())


@pytest.mark.parametrize("small_tile", [0, 128])
def test_packed_buckets_match_oracle(small_tile):
    """Ragged-mix (N 256..8192: seismic, AR(1) 0.95, gated sinusoids) and held-out
    (seismic N=2001, white noise N=8192) traces through PackedEMD's bucketed persistent
    launches; with a 128-slot LDS tile most sifts keep their knots in the global tile."""
    import torch
    from cuda_emd import EMDKernelConfig
    from cuda_emd.packed import PackedEMD
    g, sel, traces = _load([RAGGED, HELDOUT])
    lengths = np.array([len(t) for t in traces], np.int64)
    offsets = np.zeros(len(traces) + 1, np.int64)
    offsets[1:] = np.cumsum(lengths)
    cfg = EMDKernelConfig(max_extrema=1024)
    p = PackedEMD(cfg, None, _FixedTuner(small_tile)).prepare(
        torch.from_numpy(np.concatenate(traces)), offsets)
    assert {c.threads for c in p.choices.values()} == {256, 1024}
    out = p.run()
    torch.cuda.synchronize()
    gpu = _ragged_records(out, lengths)
    p.close()
    _assert_parity(f"packed(tile={small_tile})", traces, gpu, _oracle(g, sel, traces))


@pytest.mark.parametrize("threads", [256, 1024])
def test_streamed_stress_matches_oracle(threads):
    """Stress traces (N=4096: chirp mix, AR(1) 0.99, FM chirp; max_extrema 2048) streamed
    from host memory through two captured chunk graphs (5 traces per chunk, so the graphs
    replay), compacted rows compared with the oracle's."""
    from cuda_emd import EMDKernelConfig
    from cuda_emd.stream import StreamingEMD
    g, sel, traces = _load([STRESS])
    x = np.stack(traces)
    launcher = EMDKernelConfig(max_extrema=2048).build()
    s = StreamingEMD(launcher, 4096, chunk=5, threads=threads)
    rec = s.run(x, keep_rows=True)
    ro, rows = rec["row_offsets"], rec["rows"]
    n = rec["n_imfs"]
    assert np.array_equal(np.diff(ro), n.astype(np.int64) * 4096)
    gpu = dict(imfs=[rows[ro[b]:ro[b + 1]].reshape(n[b], 4096) for b in range(len(x))],
               n_imfs=n, overflow=rec["overflow"], debug_n=rec["debug_n"],
               debug_extno=rec["debug_extno"], debug_nzm=rec["debug_nzm"], stats=rec["stats"])
    _assert_parity(f"stream(threads={threads})", traces, gpu, _oracle(g, sel, traces))
