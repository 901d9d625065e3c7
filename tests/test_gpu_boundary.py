"""The boundary's guards on the GPU: queue orders are checked (host) or skipped when they
name no trace (device), and launches of one handle from two streams never share its
scratch at the same time."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(B=12, seed=61):
    from cuda_emd import synth
    rng = np.random.default_rng(seed)
    base = synth.chirp_mix(B, 1100, seed=seed)
    tr = [np.ascontiguousarray(base[k, :int(rng.integers(200, 1101))]) for k in range(B)]
    off = np.zeros(B + 1, np.int64)
    off[1:] = np.cumsum([len(t) for t in tr])
    return np.concatenate(tr), off


def _np(out):
    return [t.cpu().numpy() for t in out]


def test_host_order_must_be_a_permutation():
    import torch
    from cuda_emd import EMDKernelConfig
    v, off = _batch()
    L = EMDKernelConfig().build()
    dv = torch.from_numpy(v).cuda()
    B = len(off) - 1
    for bad in (np.r_[np.arange(B - 1), 0], np.r_[np.arange(B - 1), B], np.arange(B - 1),
                np.r_[-1, np.arange(1, B)]):
        with pytest.raises(ValueError):
            L.run_ragged(dv, off, order=bad.astype(np.int32))
    bad_off = off.copy()
    bad_off[0] = -5
    with pytest.raises(ValueError):
        L.run_ragged(dv, bad_off)


def test_device_order_entry_naming_no_trace_is_skipped():
    import torch
    from cuda_emd import EMDKernelConfig
    v, off = _batch()
    B = len(off) - 1
    L = EMDKernelConfig().build()
    dv = torch.from_numpy(v).cuda()
    ref = _np(L.run_ragged(dv, off, persistent=True, threads=256))
    for persistent in (True, False):
        order = torch.tensor(list(range(B - 1)) + [B + 100], dtype=torch.int32, device="cuda")
        # trace B-1 is named by no queue entry and the last entry names no trace: the
        # kernel skips that entry, and every named trace is decomposed as before
        got = _np(L.run_ragged(dv, off, order=order, persistent=persistent, threads=256))
        for k in range(2, 7):
            assert np.array_equal(got[k][:B - 1], ref[k][:B - 1]), (persistent, k)
        lens = np.diff(off)
        for s in range(B - 1):
            n, o = int(ref[2][s]), int(ref[1][s])
            assert np.array_equal(got[0][o:o + n * lens[s]], ref[0][o:o + n * lens[s]])


def test_one_handle_two_streams():
    """Back-to-back launches of one handle on two streams (no host sync between them)
    give the results of the same launches run one after the other."""
    import torch
    from cuda_emd import EMDKernelConfig
    va, offa = _batch(40, 62)
    vb, offb = _batch(33, 63)
    L = EMDKernelConfig().build()
    da, db = torch.from_numpy(va).cuda(), torch.from_numpy(vb).cuda()
    ref_a = _np(L.run_ragged(da, offa, persistent=True, threads=256, tile_slots=64))
    ref_b = _np(L.run_ragged(db, offb, persistent=True, threads=256, tile_slots=64))
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(torch.cuda.current_stream())
    s2.wait_stream(torch.cuda.current_stream())
    outs = []
    for rep in range(3):
        with torch.cuda.stream(s1):
            oa = L.run_ragged(da, offa, persistent=True, threads=256, tile_slots=64)
        with torch.cuda.stream(s2):
            ob = L.run_ragged(db, offb, persistent=True, threads=256, tile_slots=64)
        outs.append((oa, ob))
    torch.cuda.synchronize()
    for oa, ob in outs:
        for ref, got, off in ((ref_a, _np(oa), offa), (ref_b, _np(ob), offb)):
            for k in range(2, 8):
                assert np.array_equal(ref[k], got[k]), k
            lens = np.diff(off)
            for s in range(len(lens)):
                n, o = int(ref[2][s]), int(ref[1][s])
                assert np.array_equal(got[0][o:o + n * lens[s]], ref[0][o:o + n * lens[s]])
