"""emd_run_ragged_noalloc (queue counter and global knot tiles in a static device array,
no allocation) gives emd_run_ragged's outputs bit for bit, on two calls in a row, the
second with a smaller batch, for persistent and plain launches."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(B, n_hi, seed):
    from cuda_emd import synth
    rng = np.random.default_rng(seed)
    seis = synth.seismic_traces(B, n_hi, seed=seed)
    tr = [np.ascontiguousarray(seis[k, :int(rng.integers(64, n_hi + 1))]) for k in range(B)]
    off = np.zeros(B + 1, np.int64)
    off[1:] = np.cumsum([len(t) for t in tr])
    return np.concatenate(tr), off


def _noalloc(launcher, dv, doff, B, max_n, L):
    """Outputs allocated by the caller, then one call of the no-allocation entry point."""
    import torch
    from cuda_emd import _lib
    from cuda_emd.kernels import _ptr, _stream
    c = launcher.config
    i32 = dict(dtype=torch.int32, device="cuda")
    imfs = torch.empty((int(doff[-1].item()) * c.max_imfs,), dtype=torch.float64, device="cuda")
    out_off = doff[:-1] * c.max_imfs
    n_imfs, ovf = torch.empty((B,), **i32), torch.empty((B,), **i32)
    dn, de, dz = (torch.empty((B, c.max_imfs), **i32) for _ in range(3))
    stats = torch.empty((B, 4), **i32)
    torch.cuda.synchronize()
    _lib.check(_lib.load().emd_run_ragged_noalloc(
        launcher._h, _ptr(dv), _ptr(doff), B, max_n, None, ctypes.byref(L), _ptr(imfs),
        _ptr(out_off), _ptr(n_imfs), _ptr(ovf), _ptr(dn), _ptr(de), _ptr(dz), _ptr(stats),
        _stream()), "emd_run_ragged_noalloc")
    return imfs, out_off, n_imfs, ovf, dn, de, dz, stats


@pytest.mark.parametrize("persistent,threads,tile", [(1, 256, 0), (1, 256, 96), (0, 128, 96)])
def test_noalloc_matches_ragged(persistent, threads, tile):
    import torch
    from cuda_emd import EMDKernelConfig, _lib
    launcher = EMDKernelConfig().build()
    for B, n_hi, seed in ((600, 2001, 71), (37, 900, 72)):   # second call: smaller input
        v, off = _batch(B, n_hi, seed)
        dv, doff = torch.from_numpy(v).cuda(), torch.from_numpy(off).cuda()
        max_n = int(np.diff(off).max())
        ref = launcher.run_ragged(dv, off, threads=threads, tile_slots=tile,
                                  persistent=bool(persistent))
        ref = [t.cpu().numpy() for t in ref]
        L = _lib.EmdLaunchC(threads, tile, 0, persistent, 0)
        got = [t.cpu().numpy() for t in _noalloc(launcher, dv, doff, B, max_n, L)]
        for k in range(2, 8):
            assert np.array_equal(ref[k], got[k]), (B, k)
        if tile:
            assert np.any(got[7][:, 3] & 2)     # some sifts used the global knot tile
        lens = np.diff(off)
        for s in range(B):
            n, o = int(ref[2][s]), int(ref[1][s])
            assert np.array_equal(got[0][o:o + n * lens[s]], ref[0][o:o + n * lens[s]]), (B, s)


def test_noalloc_rejects_oversize_traces():
    """Traces longer than the kernels' 16384-sample maximum are refused before anything
    is launched (the largest workgroup region of a valid launch, 1.44 MB, fits the static
    array; larger launches run with fewer resident workgroups)."""
    import torch
    from cuda_emd import EMDKernelConfig, _lib
    launcher = EMDKernelConfig().build()
    dv = torch.zeros((16385,), dtype=torch.float64, device="cuda")
    doff = torch.tensor([0, 16385], dtype=torch.int64, device="cuda")
    L = _lib.EmdLaunchC(256, 0, 0, 1, 0)
    with pytest.raises(ValueError, match="outside"):
        _noalloc(launcher, dv, doff, 1, 16385, L)
