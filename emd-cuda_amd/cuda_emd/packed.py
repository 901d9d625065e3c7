"""packed.py -- ragged batches through length buckets, an autotuner and the cost model.

A ragged CSR batch is split by trace length into buckets (upper bounds 512, 1024, ...,
16384).  Each bucket is one persistent launch over its own queue (a subset of the batch,
addressed through the queue order), captured once into a HIP graph and replayed on its own
stream, so the tail of one bucket overlaps the next.  Per bucket:

* the cost model (when given) orders the queue longest-predicted-first (sifts x N) and
  proposes the LDS knot tile (a high quantile of predicted peak extrema);
* the autotuner picks the workgroup size (64 to 1024 lanes), the LDS tile (the capacity rule, the
  model's proposal, or the largest tile that lets 2, 3 or 4 workgroups share a CU) and the
  resident workgroups per CU by timing the candidates on a sample of the bucket's traces;
  results are cached in a JSON file.

None of these choices changes a decomposition (canonical arithmetic), which the tests
check.  No reference counterpart: the reference needs one shared N per launch and launches
one block per signal (kernels.py:1214-1220, 1343-1372).
"""
from __future__ import annotations

import ctypes
import json
import os
from dataclasses import dataclass

import numpy as np

from . import _lib
from .kernels import EMDBatchLauncher, EMDKernelConfig, _as_device_f64, _ptr, _stream

BUCKET_BOUNDS = (512, 1024, 2048, 4096, 8192, 16384)


def _torch():
    import torch
    return torch


def bucket_of(lengths: np.ndarray) -> np.ndarray:
    """Bucket index of each length: the first bound >= length."""
    return np.searchsorted(np.asarray(BUCKET_BOUNDS), np.asarray(lengths), side="left")


@dataclass
class LaunchChoice:
    threads: int
    tile_slots: int      # 0 = capacity rule
    wgs_per_cu: int      # 0 = as many as fit
    ms: float = float("nan")

    def as_dict(self):
        return {"threads": self.threads, "tile_slots": self.tile_slots,
                "wgs_per_cu": self.wgs_per_cu, "ms": self.ms}


def _device_key():
    torch = _torch()
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{getattr(p, 'gcnArchName', p.name)}:{p.multi_processor_count}"


class Autotuner:
    """Times launch candidates per length bucket on a sample of that bucket's traces."""

    def __init__(self, config: EMDKernelConfig, cache_path: str | None = None,
                 sample: int = 1024, threads=(64, 128, 256, 512, 1024)):
        self.config = config
        self.cache_path = cache_path
        self.sample = sample
        self.threads = tuple(threads)
        self.cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                self.cache = json.load(f)

    def _key(self, bucket: int, tile_hint: int) -> str:
        c = self.config
        return (f"{_device_key()}|me{c.max_extrema}|nb{c.nbsym}|it{c.max_iteration}|"
                f"imfs{c.max_imfs}|b{BUCKET_BOUNDS[bucket]}|tile{tile_hint}")

    @staticmethod
    def _largest_tile_for(launcher, threads: int, max_n: int, k: int, lo: int = 32,
                          hi: int = 8192) -> int:
        """Largest LDS tile (slots) with at least k resident workgroups per CU, or 0."""
        if launcher.max_resident(threads, lo, max_n) < k:
            return 0
        while hi - lo > 8:
            mid = (lo + hi) // 2
            if launcher.max_resident(threads, mid, max_n) >= k:
                lo = mid
            else:
                hi = mid
        return lo

    def choose(self, launcher: EMDBatchLauncher, values, offsets_host: np.ndarray,
               idx: np.ndarray, bucket: int, tile_hint: int = 0) -> LaunchChoice:
        """Best launch for the traces ``idx`` of bucket ``bucket`` (tile_hint: the cost
        model's tile, 0 = none)."""
        key = self._key(bucket, tile_hint)
        if key in self.cache:
            d = self.cache[key]
            return LaunchChoice(d["threads"], d["tile_slots"], d["wgs_per_cu"], d.get("ms", 0.0))
        torch = _torch()
        rng = np.random.default_rng(bucket)
        pick = np.sort(rng.choice(idx, min(len(idx), self.sample), replace=False))
        lens = offsets_host[pick + 1] - offsets_host[pick]
        sub_off = np.zeros(len(pick) + 1, np.int64)
        sub_off[1:] = np.cumsum(lens)
        gather = np.concatenate([np.arange(offsets_host[s], offsets_host[s + 1]) for s in pick])
        sub_vals = values[torch.from_numpy(gather).to(values.device)].contiguous()
        sub_off_t = torch.from_numpy(sub_off).cuda()
        max_n = int(lens.max())
        launcher.set_timing(True)
        best = None
        for th in self.threads:
            # LDS budgets: the capacity rule, the cost model's proposal, and the largest tile
            # that still lets k workgroups share a CU (sifts that outgrow it use the
            # workgroup's global tile)
            tiles = [0] + ([tile_hint] if tile_hint else [])
            for k in (2, 3, 4):
                t = self._largest_tile_for(launcher, th, max_n, k)
                if t and t not in tiles:
                    tiles.append(t)
            for tile in tiles:
                fit = launcher.max_resident(th, tile, max_n, persistent=True)
                for wpc in sorted({fit, max(1, fit - 1)}, reverse=True):
                    ms = []
                    for _ in range(2):
                        launcher.run_ragged(sub_vals, sub_off_t, threads=th, tile_slots=tile,
                                            wgs_per_cu=wpc, persistent=True)
                        ms.append(launcher.kernel_time_ms())
                    cand = LaunchChoice(th, tile, wpc, ms[-1])
                    if best is None or cand.ms < best.ms:
                        best = cand
        launcher.set_timing(False)
        self.cache[key] = best.as_dict()
        if self.cache_path:
            os.makedirs(os.path.dirname(os.path.abspath(self.cache_path)), exist_ok=True)
            with open(self.cache_path, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
        return best


class PackedEMD:
    """Learned-packed persistent decomposition of one ragged CSR batch.

    ``prepare`` allocates the batch's outputs, buckets the traces, runs the cost model
    (features -> MLP) for tiles and queue order, autotunes each bucket and captures one
    chunk graph per bucket.  ``run`` refreshes the queue order from the cost model and
    replays the graphs on one stream per bucket (forked from and joined back to the
    current stream).  Outputs follow run_ragged's layout (default row offsets).
    """

    def __init__(self, config: EMDKernelConfig | None = None, cost_model=None,
                 autotuner: Autotuner | None = None, tile_quantile: float = 0.995,
                 persistent: bool = True):
        """persistent=False launches one workgroup per trace in each bucket instead of the
        persistent queue (the reference's launch model applied per length bucket)."""
        self.config = config or EMDKernelConfig()
        self.cost_model = cost_model
        self.autotuner = autotuner or Autotuner(self.config)
        self.tile_quantile = tile_quantile
        self.persistent = persistent
        self.launchers = {}
        self.graphs = {}

    def prepare(self, values, offsets) -> "PackedEMD":
        torch = _torch()
        from . import costmodel
        c = self.config
        self.values = values if isinstance(values, torch.Tensor) else torch.from_numpy(
            np.ascontiguousarray(values))
        self.values = self.values.cuda().contiguous()
        off_host = np.asarray(offsets.cpu().numpy() if hasattr(offsets, "cpu") else offsets,
                              dtype=np.int64)
        self.offsets_host = off_host
        self.offsets = torch.from_numpy(off_host).cuda()
        B = len(off_host) - 1
        self.B = B
        lengths = np.diff(off_host)
        self.lengths = lengths
        dev = self.values.device
        i32 = dict(dtype=torch.int32, device=dev)
        total = int(off_host[-1])
        self.imfs = torch.empty((max(total * c.max_imfs, 1),), dtype=torch.float64, device=dev)
        self.n_imfs, self.overflow = torch.empty((B,), **i32), torch.empty((B,), **i32)
        self.debug_n, self.debug_extno, self.debug_nzm = (
            torch.empty((B, c.max_imfs), **i32) for _ in range(3))
        self.stats = torch.empty((B, 4), **i32)
        self.bucket = bucket_of(lengths)
        pred = None
        if self.cost_model is not None:
            self.feats = costmodel.features(self.values, self.offsets)
            pred = self.cost_model.predict(self.feats)
        self.choices, self.order_bufs, self.members = {}, {}, {}
        tuner_launcher = EMDBatchLauncher(None, c)
        for b in np.unique(self.bucket):
            idx = np.nonzero(self.bucket == b)[0]
            bmax = int(lengths[idx].max())
            tile_hint = 0
            if pred is not None:
                peak = np.quantile(np.maximum(pred[idx, 2], 0.0), self.tile_quantile)
                tile_hint = int(2 * np.ceil(1.1 * peak) + 4 * c.nbsym)
                tile_hint = int(np.ceil(tile_hint / 32.0) * 32)  # coarse, so the cache hits
            ch = self.autotuner.choose(tuner_launcher, self.values, off_host, idx, int(b),
                                       tile_hint)
            self.choices[int(b)] = ch
            self.members[int(b)] = idx
            self.order_bufs[int(b)] = torch.from_numpy(
                self._order(idx, pred).astype(np.int32)).to(dev)
            launcher = EMDBatchLauncher(None, c)
            self.launchers[int(b)] = launcher
            spec = _lib.EmdChunkC(
                _ptr(self.values).value, _ptr(self.offsets).value, 0, len(idx), bmax,
                _ptr(self.order_bufs[int(b)]).value,
                _lib.EmdLaunchC(ch.threads, ch.tile_slots, ch.wgs_per_cu,
                                1 if self.persistent else 0, B),
                _ptr(self.imfs).value, None, _ptr(self.n_imfs).value, _ptr(self.overflow).value,
                _ptr(self.debug_n).value, _ptr(self.debug_extno).value,
                _ptr(self.debug_nzm).value, _ptr(self.stats).value, None, None)
            torch.cuda.synchronize()
            g = ctypes.c_void_p()
            _lib.check(_lib.load().emd_chunk_graph_create(launcher._h, ctypes.byref(spec),
                                                           ctypes.byref(g)), "PackedEMD")
            self.graphs[int(b)] = g
        # buckets by predicted (or length-proxy) total work, largest first
        work = {b: float(np.sum(lengths[i] * (pred[i, 0] if pred is not None else 1.0)))
                for b, i in self.members.items()}
        self.launch_order = sorted(self.graphs, key=lambda b: -work[b])
        self.streams = {b: torch.cuda.Stream() for b in self.graphs}
        return self

    def _order(self, idx, pred):
        if pred is None:
            return idx
        sifts = np.maximum(pred[idx, 0], 1.0) if self.cost_model.orders_by_sifts else 1.0
        cost = sifts * self.lengths[idx]
        return idx[np.argsort(-cost, kind="stable")]

    def refresh_order(self) -> None:
        """Re-derive every bucket's queue order from the cost model (the learned step of
        each run: features -> MLP -> LPT order)."""
        if self.cost_model is None:
            return
        from . import costmodel
        pred = self.cost_model.predict(costmodel.features(self.values, self.offsets))
        for b, idx in self.members.items():
            self.order_bufs[b].copy_(_torch().from_numpy(self._order(idx, pred).astype(np.int32)))

    def run(self, refresh: bool = True, stream=None):
        """Decompose the prepared batch; returns run_ragged's tuple (imfs, out_offsets,
        n_imfs, overflow, debug_n, debug_extno, debug_nzm, stats).  The bucket launches
        fork from and join back to ``stream`` (default: the current stream).  The queue
        order is refreshed on a planning stream once this batch's previous run has joined,
        so two prepared batches run from two streams overlap each other's plans and
        tails."""
        torch = _torch()
        main = stream if stream is not None else torch.cuda.current_stream()
        if refresh and self.cost_model is not None:
            if getattr(self, "plan_stream", None) is None:
                self.plan_stream = torch.cuda.Stream()
            self.plan_stream.wait_stream(main)   # the order buffers are free again
            with torch.cuda.stream(self.plan_stream):
                self.refresh_order()
            main.wait_stream(self.plan_stream)
        start = torch.cuda.Event()
        start.record(main)
        for b in self.launch_order:
            s = self.streams[b]
            s.wait_event(start)
            _lib.check(_lib.load().emd_chunk_graph_launch(self.graphs[b],
                                                           ctypes.c_void_p(s.cuda_stream)),
                       "PackedEMD.run")
            main.wait_stream(s)
        out_offsets = self.offsets[:-1] * self.config.max_imfs
        return (self.imfs, out_offsets, self.n_imfs, self.overflow, self.debug_n,
                self.debug_extno, self.debug_nzm, self.stats)

    def close(self):
        for g in self.graphs.values():
            if g.value:
                _lib.load().emd_chunk_graph_destroy(g)
        self.graphs = {}

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FixedLaunch:
    """An autotuner stand-in that returns one launch shape for every bucket (for example
    the reference's: 128 lanes, the capacity-rule LDS tile)."""

    def __init__(self, threads: int = 128, tile_slots: int = 0, wgs_per_cu: int = 0):
        self.choice = LaunchChoice(threads, tile_slots, wgs_per_cu)

    def choose(self, launcher, values, offsets_host, idx, bucket, tile_hint=0):
        return self.choice


class PipelinedEMD:
    """A stream of batches through the learned-packed persistent kernel, with consecutive
    batches overlapped.

    Each ``submit`` plans its batch (features -> MLP -> queue order, and on the first batch
    the LDS tile; the FIFO order and ``tile_slots`` without a cost model) on a planning
    stream, then launches the persistent sift on the next of ``depth`` lanes.  A lane is a launcher handle (its own
    queue counter and global knot tiles) with its own stream, so a batch's sift takes the
    CU slots the previous batch's sift frees as its queue drains, instead of leaving them
    idle in that launch's tail.  The plan reads only its batch, so it also runs while the
    previous sift is still in flight.

    ``submit`` returns run_ragged's tuple without waiting; the results are ready once
    ``done(result)`` has been waited on (``torch.cuda.current_stream().wait_event``) or
    after ``torch.cuda.synchronize()``.  A batch's input must stay unchanged until then.
    Outputs are those of one run_ragged call per batch: the lanes run the same kernel.
    With ``timing``, HIP events bracket every sift launch on its lane's stream
    (``launch_ms()`` reads them).
    """

    def __init__(self, config: EMDKernelConfig | None = None, cost_model=None,
                 threads: int = 256, tile_slots: int = 0, depth: int = 2,
                 timing: bool = False):
        torch = _torch()
        if depth < 1:
            raise ValueError(f"depth must be >= 1, got {depth}")
        self.config = config or EMDKernelConfig()
        self.cost_model = cost_model
        self.threads, self.tile_slots = threads, tile_slots
        self.lanes = [(self.config.build(), torch.cuda.Stream()) for _ in range(depth)]
        self.plan_stream = torch.cuda.Stream()
        self.timing = timing
        self._events = []
        self._done = {}
        self._k = 0
        self._tile = None   # LDS tile the cost model sized on the first batch
        self.last_tile = tile_slots

    def _plan(self, values, offsets_dev, lengths_dev):
        """Queue order and LDS tile of one batch.  The first batch's plan sizes the tile
        from the predicted peak extrema (one host read); later batches keep that tile and
        compute only the order, on the device, so no step waits for its plan on the host.
        When the model's sift prediction does not rank traces the order is by length
        alone and needs no features."""
        from . import costmodel
        cm = self.cost_model
        if cm is None:
            return None, self.tile_slots
        if self.tile_slots or self._tile is not None:
            f = costmodel.features(values, offsets_dev) if cm.orders_by_sifts else None
            return cm.order(f, lengths_dev), self.tile_slots or self._tile
        f = costmodel.features(values, offsets_dev)
        plan = cm.plan(f, lengths_dev, nbsym=self.config.nbsym, quantile=0.995)
        self._tile = plan.tile_slots
        return plan.order, plan.tile_slots

    def submit(self, values, offsets, offsets_dev=None, lengths_dev=None) -> tuple:
        """Queue one CSR batch (float64 device ``values``, int64 ``offsets`` [B+1]).
        ``offsets_dev`` / ``lengths_dev`` (device int64 offsets, float64 lengths) spare the
        per-call host-to-device copies when the caller already holds them."""
        torch = _torch()
        values = _as_device_f64(values, "values")   # float64 only, as run_ragged
        if offsets_dev is None:
            offsets_dev = (offsets if isinstance(offsets, torch.Tensor)
                           else torch.from_numpy(np.asarray(offsets, np.int64))).cuda()
        if lengths_dev is None:
            lengths_dev = torch.diff(offsets_dev).double()
        caller = torch.cuda.current_stream()
        self.plan_stream.wait_stream(caller)          # the caller's writes of the input
        with torch.cuda.stream(self.plan_stream):
            order, tile = self._plan(values, offsets_dev, lengths_dev)
        self.last_tile = tile
        launcher, st = self.lanes[self._k % len(self.lanes)]
        self._k += 1
        st.wait_stream(self.plan_stream)
        st.wait_stream(caller)
        if order is not None:
            order.record_stream(st)
        with torch.cuda.stream(st):
            if self.timing:
                e0 = torch.cuda.Event(enable_timing=True)
                e0.record(st)
            out = launcher.run_ragged(values, offsets, order=order, threads=self.threads,
                                      tile_slots=tile, persistent=True)
            done = torch.cuda.Event(enable_timing=self.timing)
            done.record(st)
            if self.timing:
                self._events.append((e0, done))
        for t in (values, offsets_dev, lengths_dev):
            if isinstance(t, torch.Tensor) and t.is_cuda:
                t.record_stream(st)
                t.record_stream(self.plan_stream)
        self._done[id(out[0])] = done
        if len(self._done) > 4 * len(self.lanes):
            self._done.pop(next(iter(self._done)))
        return out

    def done(self, result):
        """The event recorded after ``result``'s sift."""
        return self._done[id(result[0])]

    def launch_ms(self, clear: bool = True) -> list:
        """Durations (ms) of the timed sift launches so far (waits for them)."""
        ms = [a.elapsed_time(b) for a, b in self._events]
        if clear:
            self._events.clear()
        return ms
