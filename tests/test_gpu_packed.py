"""Compaction, captured chunk graphs, H2D streaming and the bucketed learned-packed runner:
each must leave every per-trace output identical to the eager one-workgroup-per-trace
launch on the same traces."""
import numpy as np
import pytest
import torch

from cuda_emd import EMDKernelConfig, costmodel, synth
from cuda_emd.kernels import ChunkGraph
from cuda_emd.packed import Autotuner, PackedEMD, PipelinedEMD
from cuda_emd.stream import StreamingEMD

pytestmark = pytest.mark.gpu


def _eager(launcher, values, offsets, persistent=False):
    return launcher.run_ragged(torch.from_numpy(values).cuda(), torch.from_numpy(offsets),
                               threads=128, persistent=persistent)


def _assert_same(a, b, offsets, what=""):
    """a, b: run_ragged-style tuples over the same CSR batch."""
    for k in range(2, 7):
        assert np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy()), (what, k)
    assert np.array_equal(a[7].cpu().numpy()[:, :3], b[7].cpu().numpy()[:, :3]), what
    n = a[2].cpu().numpy()
    ia, ib = a[0].cpu().numpy(), b[0].cpu().numpy()
    oa, ob = a[1].cpu().numpy(), b[1].cpu().numpy()
    L = np.diff(offsets)
    for s in range(len(L)):
        ka = ia[oa[s]:oa[s] + n[s] * L[s]]
        kb = ib[ob[s]:ob[s] + n[s] * L[s]]
        assert np.array_equal(ka, kb), (what, s)


def test_compaction_prefix_sums_and_rows():
    launcher = EMDKernelConfig().build()
    x = synth.chirp_mix(40, 777, seed=50)
    imfs, n_imfs, *_ = launcher.run(torch.from_numpy(x).cuda())
    row, comp = launcher.compact(imfs, n_imfs)
    n = n_imfs.cpu().numpy()
    row, comp, imfs = row.cpu().numpy(), comp.cpu().numpy(), imfs.cpu().numpy()
    assert row[0] == 0 and np.array_equal(np.diff(row), n * 777)
    for s in range(40):
        assert np.array_equal(comp[row[s]:row[s + 1]].reshape(n[s], 777), imfs[s, :n[s]])


def test_chunk_graph_replays_match_eager():
    launcher = EMDKernelConfig().build()
    rng = np.random.default_rng(51)
    chunks = []
    for seed, B in ((52, 48), (53, 31)):
        base = synth.chirp_mix(B, 1500, seed=seed)
        tr = [base[k, :int(rng.integers(100, 1501))] for k in range(B)]
        off = np.zeros(B + 1, np.int64)
        off[1:] = np.cumsum([len(t) for t in tr])
        chunks.append((np.concatenate(tr), off))
    cap_vals = max(int(o[-1]) for _, o in chunks)
    g = ChunkGraph(launcher, 48, cap_vals, 1500, threads=256, tile_slots=96, compact=True)
    for rep, (v, off) in enumerate((chunks[0], chunks[1], chunks[0], chunks[0])):
        g.load(torch.from_numpy(v), torch.from_numpy(off))
        g.launch()
        torch.cuda.synchronize()
        ref = _eager(launcher, v, off)  # eager work between replays
        B = len(off) - 1
        got = (g.imfs, g.offsets[:B] * launcher.config.max_imfs, g.n_imfs[:B], g.overflow[:B],
               g.debug_n[:B], g.debug_extno[:B], g.debug_nzm[:B], g.stats[:B])
        _assert_same(ref, got, off, f"replay {rep}")
        row, comp = launcher.compact(ref[0], ref[2], offsets=off, out_offsets=ref[1])
        B1 = B + 1
        assert np.array_equal(g.row_offsets[:B1].cpu().numpy(), row.cpu().numpy())
        tot = int(row[-1])
        assert np.array_equal(g.compact_out[:tot].cpu().numpy(), comp.cpu().numpy())


def test_streaming_matches_one_shot():
    launcher = EMDKernelConfig().build()
    x = synth.chirp_mix(70, 1024, seed=54)
    s = StreamingEMD(launcher, 1024, chunk=24, threads=256)
    rec = s.run(x, keep_rows=True)
    imfs, n_imfs, ovf, dn, de, dz = launcher.run(torch.from_numpy(x).cuda())
    assert np.array_equal(rec["n_imfs"], n_imfs.cpu().numpy())
    assert np.array_equal(rec["overflow"], ovf.cpu().numpy())
    for k, t in (("debug_n", dn), ("debug_extno", de), ("debug_nzm", dz)):
        assert np.array_equal(rec[k], t.cpu().numpy())
    row, comp = launcher.compact(imfs, n_imfs)
    assert np.array_equal(rec["row_offsets"], row.cpu().numpy())
    assert np.array_equal(rec["rows"], comp.cpu().numpy())


@pytest.mark.parametrize("learned", [False, True])
def test_packed_ragged_matches_plain(tmp_path, learned):
    values, offsets, _ = synth.ragged_mix(700, seed=55, n_lo=64, n_hi=5000)
    cfg = EMDKernelConfig()
    launcher = cfg.build()
    ref = _eager(launcher, values, offsets)
    cm = None
    if learned:
        tv, to, _ = synth.ragged_mix(600, seed=56, n_lo=64, n_hi=5000)
        tr = _eager(launcher, tv, to, persistent=True)
        cm = costmodel.CostModel(seed=2)
        cm.fit(costmodel.features(torch.from_numpy(tv).cuda(), torch.from_numpy(to).cuda()),
               costmodel.targets_from(tr[7], tr[2]), epochs=10, batch=128)
    tuner = Autotuner(cfg, cache_path=str(tmp_path / "tune.json"), sample=128)
    p = PackedEMD(cfg, cost_model=cm, autotuner=tuner).prepare(values, offsets)
    for rep in range(2):
        got = p.run()
        torch.cuda.synchronize()
        _assert_same(ref, got, offsets, f"packed rep {rep}")
        _eager(launcher, values[:offsets[1]], offsets[:2])  # eager work between replays
    assert (tmp_path / "tune.json").exists()
    p.close()


@pytest.mark.parametrize("learned", [False, True])
def test_pipelined_batches_match_eager(learned):
    """Consecutive batches overlapped on two lanes (handle + stream each): every batch's
    outputs equal one eager launch over it, also when two batches' sifts run at once."""
    cfg = EMDKernelConfig()
    launcher = cfg.build()
    batches = [synth.ragged_mix(300, seed=60 + i, n_lo=64, n_hi=3000)[:2] for i in range(3)]
    refs = [_eager(launcher, v, o) for v, o in batches]
    cm = None
    if learned:
        tv, to, _ = synth.ragged_mix(400, seed=59, n_lo=64, n_hi=3000)
        tr = _eager(launcher, tv, to, persistent=True)
        cm = costmodel.CostModel(seed=3)
        cm.fit(costmodel.features(torch.from_numpy(tv).cuda(), torch.from_numpy(to).cuda()),
               costmodel.targets_from(tr[7], tr[2]), epochs=5, batch=128)
    pipe = PipelinedEMD(cfg, cost_model=cm, depth=2, timing=True)
    dev = [(torch.from_numpy(v).cuda(), torch.from_numpy(o)) for v, o in batches]
    print(f"pipelined learned={learned}", flush=True)
    outs = []
    for rnd in range(2):
        for i, (v, o) in enumerate(dev):
            outs.append((i, pipe.submit(v, o)))
    torch.cuda.synchronize()
    assert len(pipe.launch_ms()) == 6
    for i, got in outs:
        _assert_same(refs[i], got, batches[i][1], f"pipelined batch {i}")


def test_packed_instances_overlapped_on_two_streams(tmp_path):
    """Two prepared instances of a batch run alternately from two streams (plans on their
    own streams): every run's outputs equal the eager launch's."""
    values, offsets, _ = synth.ragged_mix(500, seed=70, n_lo=64, n_hi=4000)
    cfg = EMDKernelConfig()
    launcher = cfg.build()
    ref = _eager(launcher, values, offsets)
    tv, to, _ = synth.ragged_mix(400, seed=71, n_lo=64, n_hi=4000)
    tr = _eager(launcher, tv, to, persistent=True)
    cm = costmodel.CostModel(seed=4)
    cm.fit(costmodel.features(torch.from_numpy(tv).cuda(), torch.from_numpy(to).cuda()),
           costmodel.targets_from(tr[7], tr[2]), epochs=5, batch=128)
    tuner = Autotuner(cfg, cache_path=str(tmp_path / "tune.json"), sample=128)
    inst = [PackedEMD(cfg, cost_model=cm, autotuner=tuner).prepare(values, offsets)
            for _ in range(2)]
    homes = [torch.cuda.Stream(), torch.cuda.Stream()]
    print("packed instances on two streams", flush=True)
    for rnd in range(2):
        outs = [inst[i].run(stream=homes[i]) for i in range(2)]
        torch.cuda.synchronize()
        for i, got in enumerate(outs):
            _assert_same(ref, got, offsets, f"instance {i} round {rnd}")
    for p in inst:
        p.close()
