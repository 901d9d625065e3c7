"""stream.py -- host-resident batches decomposed in captured chunks with H2D streaming.

The batch stays in host memory; chunks of ``chunk`` traces are staged through pinned
buffers, copied to the device on a copy stream, and decomposed by a captured ChunkGraph
(persistent sift + compaction) on a compute stream, with two buffer slots so the copy of
chunk i+1 overlaps the decomposition of chunk i.  Per-trace records (n_imfs, overflow,
debug, stats) come back for every trace; compacted IMF rows come back only when asked for
(``keep_rows``), because a full stress batch produces far more rows than its input.
No reference counterpart: the reference decomposes one device-resident batch per call
(kernels.py:1343-1372).
"""
from __future__ import annotations

import numpy as np

from .kernels import ChunkGraph, EMDBatchLauncher


def _torch():
    import torch
    return torch


class StreamingEMD:
    def __init__(self, launcher: EMDBatchLauncher, N: int, chunk: int = 16384,
                 threads: int = 256, tile_slots: int = 0, wgs_per_cu: int = 0, slots: int = 2):
        torch = _torch()
        self.launcher, self.N, self.chunk = launcher, N, chunk
        self.graphs = [ChunkGraph(launcher, chunk, chunk * N, N, threads=threads,
                                  tile_slots=tile_slots, wgs_per_cu=wgs_per_cu,
                                  persistent=True, compact=True) for _ in range(slots)]
        self.pinned = [torch.empty((chunk * N,), dtype=torch.float64, pin_memory=True)
                       for _ in range(slots)]
        self.copy_stream = torch.cuda.Stream()
        self.compute_stream = torch.cuda.Stream()
        off = np.arange(0, (chunk + 1) * N, N, dtype=np.int64)
        self.offsets_host = torch.from_numpy(off).pin_memory()

    def run(self, signals: np.ndarray, keep_rows: bool = False) -> dict:
        """Decompose host ``signals`` [B,N] float64.  Returns per-trace numpy records and,
        with keep_rows, (row_offsets [B+1], rows) of the compacted output."""
        torch = _torch()
        x = np.asarray(signals)
        if x.dtype != np.float64 or x.ndim != 2 or x.shape[1] != self.N:
            raise ValueError(f"signals must be float64 [B, {self.N}]")
        B, N = x.shape
        c = self.launcher.config
        rec = {k: np.empty((B,), np.int32) for k in ("n_imfs", "overflow")}
        for k in ("debug_n", "debug_extno", "debug_nzm"):
            rec[k] = np.empty((B, c.max_imfs), np.int32)
        rec["stats"] = np.empty((B, 4), np.int32)
        chunk_rows = {}  # chunk index -> (local row offsets [count+1], rows)
        nslot = len(self.graphs)
        loaded = [torch.cuda.Event() for _ in range(nslot)]
        done = [torch.cuda.Event() for _ in range(nslot)]
        pending = [None] * nslot  # (chunk index, start, count, record copies) per slot

        def drain(slot):
            if pending[slot] is None:
                return
            done[slot].synchronize()
            i, start, count, dev = pending[slot]
            for k, t in dev.items():
                if k in rec:
                    rec[k][start:start + count] = t.numpy()[:count]
            if keep_rows:
                ro = dev["row_offsets"].numpy()[:count + 1].copy()
                chunk_rows[i] = (ro, self.graphs[slot].compact_out[:int(ro[count])].cpu().numpy())
            pending[slot] = None

        for i, start in enumerate(range(0, B, self.chunk)):
            slot = i % nslot
            drain(slot)  # the slot's previous chunk is finished and read back
            count = min(self.chunk, B - start)
            g = self.graphs[slot]
            self.pinned[slot][:count * N].copy_(torch.from_numpy(x[start:start + count].ravel()))
            with torch.cuda.stream(self.copy_stream):
                g.load(self.pinned[slot][:count * N], self.offsets_host[:count + 1])
                loaded[slot].record(self.copy_stream)
            self.compute_stream.wait_event(loaded[slot])
            with torch.cuda.stream(self.compute_stream):
                g.launch()
                dev = {k: getattr(g, k).to("cpu", non_blocking=True)
                       for k in ("n_imfs", "overflow", "debug_n", "debug_extno", "debug_nzm",
                                 "stats", "row_offsets")}
                done[slot].record(self.compute_stream)
            # no wait here: this slot's staging buffer is refilled only after drain(slot),
            # which waits for the chunk's decomposition, and that waited for its copy; the
            # next chunk's host staging overlaps this chunk's copy
            pending[slot] = (i, start, count, dev)
        for slot in range(nslot):
            drain(slot)
        torch.cuda.current_stream().wait_stream(self.compute_stream)
        if keep_rows:
            offs, rows, base = [np.zeros(1, np.int64)], [], 0
            for i in sorted(chunk_rows):
                ro, r = chunk_rows[i]
                offs.append(ro[1:] + base)
                rows.append(r)
                base += int(ro[-1])
            rec["row_offsets"] = np.concatenate(offs)
            rec["rows"] = np.concatenate(rows) if rows else np.zeros(0)
        return rec
