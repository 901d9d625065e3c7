#!/usr/bin/env python3
"""Benchmark: batched EMD (PyEMD EMD() defaults, fp64) traces/sec on one MI355X.

A "step" decomposes the whole synthetic seismic workload once: 16384 traces x N=2001
(25 Hz Ricker over sparse reflectivity + 5-60 Hz noise at 10 dB, seed 1), inputs already
resident in HBM.  Measured paths:

* plain:  one workgroup per trace (the reference's launch model, LDS sized by
  max_extrema), through EMDBatchLauncher.run;
* packed: the persistent queue kernel with the learned cost model ordering the queue
  (longest predicted first) and sizing the LDS tile; FIFO order when no model is given.

Prints ONE JSON line (rank 0).  ``roofline`` prices the dominant sift kernel against the
MI355X FP64 vector peak (78.6 TFLOP/s): algorithmic flops per launch (57*N per sift +
24 per knot + 2*N per output row, from exact per-trace counters the kernel records)
divided by the kernel's average duration from HIP events on its stream.  ``cpu_baseline``
times the NumPy/SciPy restatement of PyEMD EMD() (the oracle) on a bounded sample of
the same traces with a process pool on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "emd-cuda_amd"))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6
HBM_PEAK_GBS = 8000.0
METRIC = "traces/sec fp64 N=2001, learned-packed persistent vs plain kernel, 1 GPU"


def algorithmic_flops(stats: np.ndarray, n_imfs: np.ndarray, lengths: np.ndarray) -> float:
    sifts = stats[:, 0].astype(np.float64)
    knots = stats[:, 2].astype(np.float64)
    L = lengths.astype(np.float64)
    return float(np.sum(57.0 * L * sifts + 24.0 * knots + 2.0 * L * n_imfs))


def algorithmic_bytes(n_imfs: np.ndarray, lengths: np.ndarray) -> float:
    L = lengths.astype(np.float64)
    return float(np.sum(8.0 * L + 8.0 * L * n_imfs + 4 * (2 + 3 * 16 + 4)))


def profile_values(name="emd_sift_kernels_seismic_16384x2001.md") -> dict:
    """The ``key: number`` lines of a committed rocprofv3 summary under profiles/
    (counters cannot be read inside the timed run)."""
    import re
    try:
        text = open(os.path.join(ROOT, "profiles", name)).read()
    except OSError:
        return {}
    return {k: float(v) for k, v in re.findall(r"^([a-z0-9_]+):\s*([0-9.]+)\s*$", text, re.M)}


def measured_traffic(name="emd_sift_kernels_seismic_16384x2001.md"):
    """HBM bytes per launch of a kernel on its workload, from its FETCH_SIZE/WRITE_SIZE
    summary."""
    return profile_values(name).get("traffic_bytes_per_launch")


def _cpu_one(x, T=None):
    from oracle import pyemd_restated as O
    return O.emd(x, max_imf=15, T=T).n_imfs


def _cpu_record(x, T=None):
    from oracle import pyemd_restated as O
    return O.emd(x, max_imf=15, T=T)


def cpu_baseline(x: np.ndarray, n: int, T=None):
    """The oracle timed on the first n traces with a process pool (the reported CPU
    baseline), then run again untimed to return its full records for the parity check of
    the GPU output on the same traces.  T: a time axis shared by the traces (the
    reference's T), or None."""
    import functools
    import multiprocessing as mp
    cores = min(16, os.cpu_count() or 1)
    sample = x[:n]
    one = functools.partial(_cpu_one, T=T)
    rec = functools.partial(_cpu_record, T=T)
    ctx = mp.get_context("fork")
    with ctx.Pool(cores) as pool:
        pool.map(one, sample[:cores])  # warm the workers (imports)
        t0 = time.perf_counter()
        pool.map(one, sample, chunksize=max(1, n // (4 * cores)))
        dt = time.perf_counter() - t0
        records = pool.map(rec, sample, chunksize=max(1, n // (4 * cores)))
    axis = "" if T is None else " on a shared jittered time axis"
    return {"value": n / dt, "unit": "traces/s", "cores": cores, "kind": "port",
            "sample": f"first {n} of the {len(x)} seismic traces (N={x.shape[1]}){axis}, "
                      f"NumPy/SciPy PyEMD EMD() restatement, multiprocessing pool of {cores}"}, records


def ragged_records(out, lengths, count=None):
    """run_ragged's tuple as the per-trace dict tests/parity.py compares (first count
    traces)."""
    count = len(lengths) if count is None else count
    imfs, oo = out[0].cpu().numpy(), out[1].cpu().numpy()
    n = out[2].cpu().numpy()
    rows = [imfs[oo[s]:oo[s] + n[s] * lengths[s]].reshape(n[s], lengths[s])
            for s in range(count)]
    return dict(imfs=rows, n_imfs=n[:count], overflow=out[3].cpu().numpy()[:count],
                debug_n=out[4].cpu().numpy()[:count], debug_extno=out[5].cpu().numpy()[:count],
                debug_nzm=out[6].cpu().numpy()[:count], stats=out[7].cpu().numpy()[:count])


def parity_vs_oracle(x, out, lengths, records):
    """GPU output of the timed path against the oracle on the CPU-baseline traces."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from parity import compare
    n = len(records)
    bad, worst = compare(list(x[:n]), ragged_records(out, lengths, n), records,
                         label="bench:seismic")
    traces_bad = sorted({m[0] for m in bad})
    unexplained = sorted({m[0] for m in bad if not m[2]})
    return {"traces": n, "mismatched_traces": len(traces_bad),
            "tie_explained": len(traces_bad) - len(unexplained),
            "unexplained": len(unexplained), "worst_imf_err_rel": worst,
            "bar": "n_imfs, overflow, debug_n/extno/nzm and sift/peak/knot counters "
                   "identical; IMF max-abs <= 1e-9 max|x|"}


def train_cost_model(launcher, args):
    """Train the cost model (untimed) on the method's own counts for seismic traces of a
    different seed (3) than the timed workload (1)."""
    import torch
    from cuda_emd import costmodel, synth
    xt = synth.seismic_traces(args.train, args.n, seed=3)
    v = torch.from_numpy(np.ascontiguousarray(xt.ravel())).cuda()
    off = torch.arange(0, xt.size + 1, xt.shape[1], dtype=torch.int64).cuda()
    t0 = time.perf_counter()
    out = launcher.run_ragged(v, off, threads=args.threads, persistent=True)
    f = costmodel.features(v, off)
    cm = costmodel.CostModel(seed=0)
    hist = cm.fit(f, costmodel.targets_from(out[7], out[2]), epochs=args.epochs, batch=1024)
    torch.cuda.synchronize()
    return cm, {"train_traces": int(args.train), "train_seed": 3, "epochs": args.epochs,
                "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                "train_s": time.perf_counter() - t0,
                "sift_rank_holdout": cm.sift_rank,
                "queue_order": "predicted sifts x N" if cm.orders_by_sifts else
                               "N, stable (predicted sifts do not rank held-out traces)"}


def bench_ragged(args):
    """Ragged mix: plain one-workgroup-per-trace kernel over the CSR batch, versus the
    bucketed persistent runner in FIFO order, versus the same runner ordered and
    LDS-sized by the cost model (trained on a seed-3 ragged mix)."""
    import torch
    from cuda_emd import EMDKernelConfig, costmodel, synth
    from cuda_emd.packed import Autotuner, PackedEMD
    B = args.batch if args.batch != 16384 else 65536
    values, offsets, kinds = synth.ragged_mix(B, seed=2)
    lengths = np.diff(offsets)
    cfg = EMDKernelConfig.for_device()
    launcher = cfg.build()
    dv = torch.from_numpy(values).cuda()
    doff = torch.from_numpy(offsets).cuda()

    def timed(fn):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        torch.cuda.synchronize()
        return out, (time.perf_counter() - t0) / args.steps * 1e3

    res = {}
    # plain: the reference's launch model (one 128-lane workgroup per trace, the
    # capacity-rule LDS tile) applied per length bucket, each bucket's launch captured and
    # replayed on its own stream like the packed runner's
    from cuda_emd.packed import FixedLaunch
    plain = PackedEMD(cfg, None, FixedLaunch(128, 0, 0), persistent=False).prepare(dv, offsets)
    res["plain"] = timed(lambda: plain.run())
    _progress(f"plain per bucket: {res['plain'][1]:.1f} ms")
    single = timed(lambda: launcher.run_ragged(dv, doff, persistent=False))[1]
    _progress(f"plain single launch: {single:.1f} ms")
    tuner = Autotuner(cfg, cache_path=args.tune_cache)
    fifo = PackedEMD(cfg, None, tuner).prepare(dv, offsets)
    res["fifo"] = timed(lambda: fifo.run())
    tv, to, _ = synth.ragged_mix(min(args.train, 16384), seed=3)
    tvd, tod = torch.from_numpy(tv).cuda(), torch.from_numpy(to).cuda()
    tr = launcher.run_ragged(tvd, tod, threads=256, persistent=True)
    cm = costmodel.CostModel(seed=0)
    hist = cm.fit(costmodel.features(tvd, tod), costmodel.targets_from(tr[7], tr[2]),
                  epochs=args.epochs, batch=1024)
    learned = PackedEMD(cfg, cm, tuner).prepare(dv, offsets)
    res["packed"] = timed(lambda: learned.run())
    a = res["plain"][0]
    ident = {}
    for name in ("fifo", "packed"):
        b = res[name][0]
        ident[name] = bool(all(np.array_equal(a[k].cpu().numpy(), b[k].cpu().numpy())
                               for k in range(2, 7)))
    stats, n_imfs = res["packed"][0][7].cpu().numpy(), res["packed"][0][2].cpu().numpy()
    flops = algorithmic_flops(stats, n_imfs, lengths)
    ms = res["packed"][1]
    achieved = flops / (ms * 1e-3) / 1e12
    pred = cm.predict(costmodel.features(dv, doff))
    true = costmodel.targets_from(res["packed"][0][7], res["packed"][0][2])
    line = {
        "metric": "traces/sec fp64 ragged N in [256,8192], learned-packed persistent vs plain "
                  "kernel, 1 GPU",
        "value": B / (ms * 1e-3), "unit": "traces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"ragged mix, {B} traces, N log-uniform [256, 8192], seed 2",
                   "traces": B, "mean_N": float(lengths.mean())},
        # the compute roof: MI355X's fp64 peak (78.6 TFLOP/s, the same for its vector ALUs,
        # which run the sift, as for its matrix cores, which it does not use)
        "roofline": {"bound": "valu", "roof": "fp64 VALU (78.6 TFLOP/s; the sift issues no MFMA)", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_step": flops,
                     "note": "step time of all bucket launches (concurrent streams)"},
        "plain_traces_per_s": B / (res["plain"][1] * 1e-3),
        "plain": "one 128-lane workgroup per trace, capacity-rule LDS tile, one launch per "
                 "length bucket (buckets on their own streams)",
        "plain_single_launch_traces_per_s": B / (single * 1e-3),
        "fifo_traces_per_s": B / (res["fifo"][1] * 1e-3),
        "speedup_vs_plain": res["plain"][1] / ms,
        "speedup_vs_fifo": res["fifo"][1] / ms,
        "identical_to_plain": ident,
        "buckets": {str(b): learned.choices[b].as_dict() | {"traces": int(len(learned.members[b]))}
                    for b in sorted(learned.choices)},
        "cost_model": {"loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
                       "mape_on_workload": costmodel.mape(pred, true)},
        "overflow_traces": int((res["packed"][0][3].cpu().numpy() != 0).sum()),
    }
    print(json.dumps(line), flush=True)


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _seismic_chunks(total, seed, chunk=16384):
    """total seismic traces (N=2001) drawn chunk by chunk from one generator."""
    from cuda_emd import synth
    return synth.seismic_chunks(total, seed, chunk=chunk, progress=_progress)


def bench_costmodel(args):
    """Cost model at full size: trained on the method's own counts for --train-total seismic
    traces (seed 3), evaluated (MAPE) on --eval-total held-out traces (seed 4): seismic plus
    5% N(0,1) white noise at N=8192.  Also times the held-out batch through the bucketed
    persistent runner in FIFO order and in learned order/tiles."""
    import torch
    from cuda_emd import EMDKernelConfig, costmodel
    from cuda_emd.packed import Autotuner, PackedEMD
    cfg = EMDKernelConfig.for_device()
    launcher = cfg.build()
    t0 = time.perf_counter()
    xtr = _seismic_chunks(args.train_total, 3)
    ftr, ttr = [], []
    for s in range(0, len(xtr), 32768):
        x = torch.from_numpy(xtr[s:s + 32768]).cuda()
        off = torch.arange(0, x.numel() + 1, 2001, dtype=torch.int64).cuda()
        out = launcher.run_ragged(x.view(-1), off, threads=256, tile_slots=800, persistent=True)
        ftr.append(costmodel.features(x.view(-1), off))
        ttr.append(costmodel.targets_from(out[7], out[2]))
    ftr, ttr = torch.cat(ftr), np.concatenate(ttr)
    gen_train_s = time.perf_counter() - t0
    _progress("training set decomposed")
    cm = costmodel.CostModel(seed=0)
    t1 = time.perf_counter()
    hist = cm.fit(ftr, ttr, epochs=args.epochs, batch=4096)
    torch.cuda.synchronize()
    fit_s = time.perf_counter() - t1
    _progress(f"fit done in {fit_s:.1f}s")
    # held-out: 95% seismic (N=2001) + 5% white N(0,1) at N=8192, shuffled
    from cuda_emd import synth
    values, offsets, kinds = synth.heldout_mix(args.eval_total, seed=4, progress=_progress)
    n_white = int((kinds == 1).sum())
    n_seis = len(kinds) - n_white
    dv, doff = torch.from_numpy(values).cuda(), torch.from_numpy(offsets).cuda()
    tuner = Autotuner(cfg, cache_path=args.tune_cache)
    res = {}
    # plain: the reference's launch model (one 128-lane workgroup per trace, capacity-rule
    # tile) applied per length bucket, buckets on their own streams
    from cuda_emd.packed import FixedLaunch
    pp = PackedEMD(cfg, None, FixedLaunch(128, 0, 0), persistent=False).prepare(dv, offsets)
    pp.run()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    pout = pp.run()
    torch.cuda.synchronize()
    res["plain"] = (time.perf_counter() - ts, [pout[k].cpu().numpy() for k in range(8)[2:]])
    _progress(f"held-out plain: {res['plain'][0]:.2f}s")
    pp.close()
    del pout, pp
    torch.cuda.empty_cache()
    for name, model in (("fifo", None), ("learned", cm)):
        p = PackedEMD(cfg, model, tuner).prepare(dv, offsets)
        p.run()
        torch.cuda.synchronize()
        ts = time.perf_counter()
        out = p.run()
        torch.cuda.synchronize()
        res[name] = (time.perf_counter() - ts, [out[k].cpu().numpy() for k in range(8)[2:]])
        _progress(f"held-out {name}: {res[name][0]:.2f}s")
        p.close()
        del p, out
        torch.cuda.empty_cache()
    out = res["learned"][1]  # n_imfs, overflow, debug_n, debug_extno, debug_nzm, stats
    true = costmodel.targets_from(out[5], out[0])
    pred = cm.predict(costmodel.features(dv, doff))
    same = all(np.array_equal(a, b) for a, b in zip(res["fifo"][1][:5], out[:5])) and \
        all(np.array_equal(a, b) for a, b in zip(res["plain"][1][:5], out[:5]))
    B = len(kinds)
    line = {
        "metric": "cost-model MAPE on held-out traces (sifts, n_imfs, peak extrema)",
        "value": costmodel.mape(pred, true)["sifts"], "unit": "MAPE (sifts)",
        "higher_is_better": False, "n_gpus": 1, "dtype": "f32 MLP / f64 EMD",
        "data": "synthetic",
        "config": {"workload": f"train {args.train_total} seismic (seed 3); eval {B} = "
                               f"{n_seis} seismic (seed 4) + {n_white} white N=8192"},
        "mape_all": costmodel.mape(pred, true),
        "mape_seismic": costmodel.mape(pred[kinds == 0], true[kinds == 0]),
        "mape_white": costmodel.mape(pred[kinds == 1], true[kinds == 1]),
        "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1], "epochs": args.epochs,
        "train_decompose_and_features_s": gen_train_s, "fit_s": fit_s,
        "eval_plain_traces_per_s": B / res["plain"][0],
        "eval_fifo_traces_per_s": B / res["fifo"][0],
        "eval_learned_traces_per_s": B / res["learned"][0],
        "learned_speedup_vs_fifo": res["fifo"][0] / res["learned"][0],
        "learned_speedup_vs_plain": res["plain"][0] / res["learned"][0],
        "identical_plain_fifo_learned": bool(same),
        "overflow_traces": int((out[1] != 0).sum()),
    }
    print(json.dumps(line), flush=True)


def bench_stress(args):
    """262144 x N=4096 (seed 5) with max_extrema 2048, streamed from host memory through
    captured chunks (H2D on a copy stream overlapping the persistent sift + compaction)."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    from cuda_emd.stream import StreamingEMD
    B = args.batch if args.batch != 16384 else 262144
    _progress("generating stress mix")
    x, kinds = synth.stress_mix(B, 4096, seed=5)
    _progress("generated")
    cfg = EMDKernelConfig(max_extrema=2048)
    launcher = cfg.build()
    s = StreamingEMD(launcher, 4096, chunk=16384, threads=256, tile_slots=args.tile)
    s.run(x[:16384])  # warm (first replays)
    _progress("warm")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rec = s.run(x)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    n = rec["n_imfs"].astype(np.int64)
    line = {
        "metric": "traces/sec fp64 N=4096 streamed from host (H2D included), 1 GPU",
        "value": B / dt, "unit": "traces/s", "n_gpus": 1, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic", "seconds": dt,
        "config": {"workload": f"stress mix {B} x N=4096, seed 5", "max_extrema": 2048,
                   "chunk": 16384, "tile_slots": args.tile},
        "overflow_traces": int((rec["overflow"] != 0).sum()),
        "overflow_by_kind": {k: int(((rec["overflow"] != 0) & (kinds == i)).sum())
                             for i, k in enumerate(("chirp", "ar1_0.99", "fm_chirp"))},
        "imf_cap_hits": int((rec["stats"][:, 3] & 1).sum()),
        "compacted_rows": int(n.sum()), "padded_rows": int(B * cfg.max_imfs),
        "compacted_bytes": int(n.sum() * 4096 * 8),
        "input_bytes": int(B * 4096 * 8),
        "sifts_per_trace": float(rec["stats"][:, 0].mean()),
    }
    print(json.dumps(line), flush=True)


def segy_image(B: int, N: int, seed: int) -> np.ndarray:
    """A SEG-Y file image (rev 1, IBM float samples, 4 ms) of B traces x N samples whose
    samples are random normalised IBM words of magnitude 16^-3 .. 16^2."""
    rng = np.random.default_rng(seed)
    img = np.zeros(3600 + B * (240 + 4 * N), dtype=np.uint8)
    img[3216:3218] = (4000 >> 8, 4000 & 255)          # sample interval (us)
    img[3220:3222] = (N >> 8, N & 255)                # samples per trace
    img[3225] = 1                                     # IBM float
    tr = img[3600:].reshape(B, 240 + 4 * N)
    tr[:, 114], tr[:, 115] = N >> 8, N & 255
    words = ((rng.integers(0, 2, (B, N), dtype=np.uint32) << 31)
             | (rng.integers(61, 67, (B, N), dtype=np.uint32) << 24)
             | rng.integers(0x100000, 0x1000000, (B, N), dtype=np.uint32))
    tr[:, 240:] = words.astype(">u4").view(np.uint8).reshape(B, 4 * N)
    return img


def bench_segy(args):
    """SEG-Y in: the decode kernel on a file image already in HBM (16384 traces x N=2001,
    IBM float), and the same with the pinned host -> device copy of the trace bytes."""
    import torch
    from cuda_emd import traceio
    B, N = args.batch, args.n
    img = segy_image(B, N, seed=6)
    lay = traceio.scan_segy(img)
    host = torch.from_numpy(img[lay.data_start:]).pin_memory()
    dev = host.to("cuda")
    out = torch.empty((B, N), dtype=torch.float64, device="cuda")
    for _ in range(args.warmup):
        traceio.decode_segy(dev, lay, out=out)
    torch.cuda.synchronize()
    st = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for _ in range(args.steps):
        traceio.decode_segy(dev, lay, out=out)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dev.copy_(host, non_blocking=True)
        traceio.decode_segy(dev, lay, out=out)
    torch.cuda.synchronize()
    ms_h2d = (time.perf_counter() - t0) * 1e3 / args.steps
    # The same image as a file, read chunk by chunk (memory map -> pinned staging -> H2D ->
    # decode); the file was just written, so it is read from the page cache.
    import tempfile
    with tempfile.NamedTemporaryFile(suffix=".sgy") as f:
        img.tofile(f.name)
        for _ in traceio.iter_segy(f.name, chunk=4096):
            pass
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in traceio.iter_segy(f.name, chunk=4096):
            pass
        torch.cuda.synchronize()
        ms_file = (time.perf_counter() - t0) * 1e3
    alg = 12.0 * N * B                      # 4 B read + 8 B written per sample
    achieved = alg / (ms * 1e-3) / 1e9
    line = {
        "metric": "SEG-Y traces decoded to fp64 per second (IBM float, N=2001), 1 GPU",
        "value": B / (ms * 1e-3), "unit": "traces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32->f64", "data": "synthetic",
        "config": {"workload": f"SEG-Y image, {B} traces x N={N}, IBM float, seed 6",
                   "traces": B, "N": N, "file_bytes": int(img.size)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic("segy_decode_16384x2001.md"),
                     "algorithmic_bytes": alg},
        "h2d_inclusive_traces_per_s": B / (ms_h2d * 1e-3),
        "h2d_inclusive_ms": ms_h2d,
        "file_chunked_traces_per_s": B / (ms_file * 1e-3),
        "file_chunked_ms": ms_file,
    }
    print(json.dumps(line), flush=True)


def bench_timeaxis(args):
    """The time-axis kernels (EMDKernelConfig(time_axis='given'), the reference's use of
    T): the seismic traces (16384 x N=2001, seed 1) on one jittered axis shared by the
    batch (steps of 4 ms x U(0.8, 1.2), seed 7), decomposed by the padded time-axis
    entry point (the reference's launch model, one workgroup per trace) at the packed
    path's shape (256 lanes, the largest knot tile that keeps 4 workgroups per CU), timed
    with HIP events beside the index kernel at its own packed shape on the same traces."""
    import torch
    from cuda_emd import EMDKernelConfig, synth
    x = synth.seismic_traces(args.batch, args.n, seed=1)
    B, N = x.shape
    rng = np.random.default_rng(7)
    T = 0.004 * np.cumsum(rng.uniform(0.8, 1.2, N))
    Tn = (T - T[0]) / np.min(np.diff(T))  # the reference's normalisation (emd.py:83-88)
    cpu, records = (None, None) if args.no_cpu else cpu_baseline(x, args.cpu_sample, T=T)
    dx = torch.from_numpy(x).cuda()
    dT = torch.from_numpy(Tn).cuda()
    lengths = np.full(B, N, dtype=np.int64)

    def timed(cfg, with_axis):
        L = cfg.build()
        L.set_timing(True)
        out = None
        for _ in range(args.warmup):
            out = L.run(dx, dT if with_axis else None)
        torch.cuda.synchronize()
        ms = []
        for _ in range(args.steps):
            out = L.run(dx, dT if with_axis else None)
            torch.cuda.synchronize()
            ms.append(L.kernel_time_ms())
        return out, L.last_stats, float(np.mean(ms))

    lds = 40960 - EMDKernelConfig(threads_per_block=256).static_shared_mem_bytes() \
        - 568 * ((N + 63) // 64)
    tile_t, tile_i = lds // 32, lds // 28
    out_t, st_t, ms_t = timed(EMDKernelConfig(time_axis="given", threads_per_block=256,
                                              tile_slots=tile_t), True)
    _, _, ms_i = timed(EMDKernelConfig(threads_per_block=256, tile_slots=tile_i), False)
    n_imfs = out_t[1].cpu().numpy()
    stats = st_t.cpu().numpy()
    flops = algorithmic_flops(stats, n_imfs, lengths)
    achieved = flops / (ms_t * 1e-3) / 1e12
    line = {
        "metric": "traces/sec fp64 N=2001 on a shared non-uniform time axis (the reference's "
                  "T), one workgroup per trace, 1 GPU",
        "value": B / (ms_t * 1e-3), "unit": "traces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_t, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"seismic traces, {B} x N={N}, seed 1, jittered time axis "
                               "(4 ms x U(0.8, 1.2), seed 7)", "traces": B, "N": N,
                   "threads": 256, "tile_slots": tile_t},
        "roofline": {"bound": "valu", "roof": "fp64 VALU (78.6 TFLOP/s)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": None,
                     "flops_per_launch": flops, "kernel_ms": ms_t},
        "index_axis_same_shape": {"kernel_ms": ms_i, "traces_per_s": B / (ms_i * 1e-3),
                                  "tile_slots": tile_i},
        "sifts_per_trace": float(stats[:, 0].mean()),
    }
    if cpu is not None:
        line["cpu_baseline"] = cpu
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from parity import compare
        n = len(records)
        gpu = dict(imfs=out_t[0].cpu().numpy()[:n], n_imfs=n_imfs[:n],
                   overflow=out_t[2].cpu().numpy()[:n], debug_n=out_t[3].cpu().numpy()[:n],
                   debug_extno=out_t[4].cpu().numpy()[:n],
                   debug_nzm=out_t[5].cpu().numpy()[:n], stats=stats[:n])
        bad, worst = compare(list(x[:n]), gpu, records, label="bench:timeaxis")
        traces_bad = sorted({m[0] for m in bad})
        unexplained = sorted({m[0] for m in bad if not m[2]})
        line["parity_vs_oracle"] = {"traces": n, "mismatched_traces": len(traces_bad),
                                    "tie_explained": len(traces_bad) - len(unexplained),
                                    "unexplained": len(unexplained),
                                    "worst_imf_err_rel": worst}
    print(json.dumps(line), flush=True)


def hilbert_work(lengths: np.ndarray, n_imfs: np.ndarray, n_fft: int):
    """(algorithmic bytes, algorithmic flops) of one Hilbert launch.  Bytes: each IMF sample
    is read once and its amplitude, phase and frequency written once (32 B).  Flops: per
    row pair one forward and one inverse complex transform of the row's transform length F
    (5 F log2 F each, the radix-2 count), per trace in the exact form one more forward
    transform (its multiplier), 4 per point for the multiplier product, and 12 per sample
    for the outputs' arithmetic outside atan2 and sqrt."""
    L = lengths.astype(np.int64)
    rows = n_imfs.astype(np.float64)
    samples = float(np.sum(rows * L))
    pow2 = (L & (L - 1)) == 0
    if n_fft > 0:
        F = np.full(len(L), n_fft, np.float64)
        exact = np.zeros(len(L), bool)
    elif n_fft < 0:
        F = 2.0 ** np.ceil(np.log2(np.maximum(L, 1)))
        exact = np.zeros(len(L), bool)
    else:
        exact = (~pow2) & (L > 2)
        F = np.where(exact, 2.0 ** np.ceil(np.log2(np.maximum(2 * L - 1, 1))), L.astype(float))
    fft = 5.0 * F * np.log2(np.maximum(F, 2))
    pairs = np.ceil(rows / 2)
    flops = np.sum(pairs * (2 * fft + 4 * F) + exact * (rows > 0) * fft) + 12.0 * samples
    return 32.0 * samples, float(flops)


def _hilbert_cpu_one(args):
    from oracle import hilbert_ref as H
    rows, dt = args
    for r in rows:
        H.instantaneous(r, dt=dt)
    return len(rows)


def bench_hilbert(args):
    """The Hilbert spectral step on the IMFs of the seismic workload (16384 traces x N=2001,
    seed 1): the traces are decomposed once (untimed), their IMF rows compacted back to back,
    and each timed step computes amplitude, phase and frequency of every row (dt = 4 ms)."""
    import torch
    from multiprocessing import Pool

    from cuda_emd import EMDKernelConfig, synth
    from cuda_emd.hilbert import HilbertStep
    from oracle import hilbert_ref as H

    B, N, dt = args.batch, args.n, 0.004
    x = synth.seismic_traces(B, N, seed=1)
    # the CPU-baseline workers fork now, before this process touches the GPU
    ncpu = min(16, os.cpu_count() or 1)
    pool = None if args.no_cpu else Pool(ncpu)
    launcher = EMDKernelConfig.for_device().build()
    _progress("hilbert: decomposing the input traces")
    imfs, n_imfs = launcher.run(torch.from_numpy(x).cuda())[:2]
    row, comp = launcher.compact(imfs, n_imfs)
    del imfs
    offsets = torch.arange(0, (B + 1) * N, N, dtype=torch.int64, device="cuda")
    nim = n_imfs.cpu().numpy()
    lengths = np.full(B, N, np.int64)
    res = {}
    for n_fft in (0, -1):
        step = HilbertStep(comp, n_imfs, offsets=offsets, row_base=row[:-1], dt=dt, n_fft=n_fft)
        for _ in range(args.warmup):
            step.run()
        torch.cuda.synchronize()
        st = torch.cuda.current_stream()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps + 1)]
        ev[0].record(st)
        for i in range(args.steps):
            step.run()
            ev[i + 1].record(st)
        torch.cuda.synchronize()
        per = [ev[i].elapsed_time(ev[i + 1]) for i in range(args.steps)]
        res[n_fft] = dict(ms=float(np.mean(per)), ms_min=float(np.min(per)), step=step)
        _progress(f"hilbert n_fft={n_fft}: {res[n_fft]['ms']:.3f} ms per launch")
    ms = res[0]["ms"]
    alg_bytes, alg_flops = hilbert_work(lengths, nim, 0)
    gbs = alg_bytes / (ms * 1e-3) / 1e9
    tfs = alg_flops / (ms * 1e-3) / 1e12
    # parity of the timed outputs with the oracle on a sample of traces
    out = {k: v.cpu().numpy() for k, v in res[0]["step"].out.items()}
    comp_h, rows_h = comp.cpu().numpy(), row.cpu().numpy()
    worst = {"amp": 0.0, "phase": 0.0, "freq": 0.0}
    sample = range(0, B, max(1, B // 256))
    for s in sample:
        for k in range(nim[s]):
            a = rows_h[s] + k * N
            r = comp_h[a:a + N]
            ra, rp, rf = H.instantaneous(r, dt=dt)
            well = ra >= 1e-3 * ra.max()
            sc = np.max(np.abs(r))
            worst["amp"] = max(worst["amp"], float(np.max(np.abs(out["amp"][a:a + N] - ra)) / sc))
            dp = np.angle(np.exp(1j * (out["phase"][a:a + N] - rp)))
            worst["phase"] = max(worst["phase"], float(np.max(np.abs(dp[well]), initial=0)))
            m = np.minimum(np.arange(N), N - 2)
            df = (out["freq"][a:a + N] - rf) * dt
            df -= np.round(df)
            worst["freq"] = max(worst["freq"], float(np.max(np.abs(df[well[m] & well[m + 1]]),
                                                          initial=0)))
    cpu = None
    if pool is not None:
        # the oracle (scipy.signal.hilbert + numpy) on the rows of a bounded sample of traces
        pick = list(range(0, B, max(1, B // (64 * ncpu))))
        rows = [comp_h[rows_h[s] + k * N: rows_h[s] + (k + 1) * N] for s in pick
                for k in range(nim[s])]
        chunks = [(rows[i::ncpu], dt) for i in range(ncpu)]
        with pool:
            pool.map(_hilbert_cpu_one, [(c[0][:2], dt) for c in chunks])
            t0 = time.perf_counter()
            done = sum(pool.map(_hilbert_cpu_one, chunks))
            sec = time.perf_counter() - t0
        cpu = {"value": done * N / sec, "unit": "samples/s", "cores": ncpu, "kind": "port",
               "sample": f"{done} IMF rows of {len(pick)} traces, scipy.signal.hilbert + numpy "
                         "angle/unwrap/diff, one process per core"}
    samples = float(np.sum(nim)) * N
    line = {
        "metric": "Hilbert spectral step: IMF samples per second (amplitude, phase, frequency; "
                  "fp64), seismic IMFs 16384 x N=2001, 1 GPU",
        "value": samples / (ms * 1e-3), "unit": "samples/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"IMFs of the seismic traces ({B} x N={N}, seed 1), compacted; "
                               "scipy.signal.hilbert semantics (DFT of the row's length), "
                               "dt = 4 ms", "traces": B, "N": N,
                   "imf_rows": int(np.sum(nim))},
        "roofline": {"bound": "valu" if tfs / FP64_PEAK_TFLOPS > gbs / HBM_PEAK_GBS else "hbm",
                     "achieved_gbs": gbs, "peak_gbs": HBM_PEAK_GBS, "frac_hbm": gbs / HBM_PEAK_GBS,
                     "achieved_tflops": tfs, "peak_tflops": FP64_PEAK_TFLOPS,
                     "frac_fp64": tfs / FP64_PEAK_TFLOPS, "kernel_ms": ms,
                     "algorithmic_bytes": alg_bytes, "algorithmic_flops": alg_flops,
                     "traffic": measured_traffic("emd_hilbert_seismic_imfs_16384x2001.md")},
        "traces_per_s": B / (ms * 1e-3),
        "pow2_padded": {"n_fft": -1, "ms_per_step": res[-1]["ms"],
                        "samples_per_s": samples / (res[-1]["ms"] * 1e-3)},
        "parity_vs_oracle": {"traces": len(sample), "max_amp_err_rel": worst["amp"],
                             "max_phase_err_rad": worst["phase"],
                             "max_freq_err_cycles_per_sample": worst["freq"]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--n", type=int, default=2001)
    ap.add_argument("--mode", choices=["both", "plain", "fifo", "packed", "copy"],
                    default="both",
                    help="plain: one workgroup per trace; fifo: persistent queue in batch "
                         "order with the capacity-rule tile; packed: persistent queue ordered "
                         "and LDS-sized by the cost model; both: all three; copy: only a 1 GiB "
                         "device copy, the size calibration for FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--cpu-sample", type=int, default=512)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--tile", type=int, default=0,
                    help="LDS knot slots of the packed path (0: sized by the cost model)")
    ap.add_argument("--threads", type=int, default=256)
    ap.add_argument("--lanes", type=int, default=2,
                    help="packed path: handles/streams consecutive steps rotate over")
    ap.add_argument("--serial-steps", action="store_true",
                    help="packed path: each step's sift starts after the previous one ends "
                         "(default: consecutive steps overlap on two handles and streams)")
    ap.add_argument("--serial-plan", action="store_true",
                    help="packed path: compute each step's plan on the launch stream, after "
                         "the previous step's sift (default: on a planning stream)")
    ap.add_argument("--train", type=int, default=8192,
                    help="seismic traces (seed 3) the cost model is trained on before timing")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--workload", choices=["seismic", "ragged", "costmodel", "stress", "segy",
                                           "hilbert", "timeaxis"],
                    default="seismic",
                    help="seismic: the headline 16384 x N=2001 traces; ragged: the ragged mix "
                         "(N log-uniform in [256, 8192], seed 2; --batch sets B, default 65536); "
                         "costmodel: train on 200k seismic decompositions (seed 3), evaluate "
                         "on 262144 held-out traces (seed 4, 5%% white noise at N=8192); "
                         "stress: 262144 x N=4096 (seed 5) streamed from host in captured "
                         "chunks, max_extrema 2048, compacted output; segy: decode of a "
                         "16384 x N=2001 IBM-float SEG-Y image in HBM; hilbert: amplitude, "
                         "phase and frequency of the seismic traces' compacted IMFs; "
                         "timeaxis: the seismic traces on a shared jittered time axis "
                         "through the time-axis kernels (the reference's T)")
    ap.add_argument("--train-total", type=int, default=200000)
    ap.add_argument("--eval-total", type=int, default=262144)
    ap.add_argument("--tune-cache", default=None, help="autotuner JSON cache (ragged)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); every trace is "
              "independent, so N GPUs would shard the batch with no exchange.", file=sys.stderr)
        sys.exit(2)

    if args.mode == "copy":
        import torch
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": src.numel() * 8, "launches": 3}), flush=True)
        return

    if args.workload == "ragged":
        return bench_ragged(args)
    if args.workload == "costmodel":
        return bench_costmodel(args)
    if args.workload == "stress":
        return bench_stress(args)
    if args.workload == "segy":
        return bench_segy(args)
    if args.workload == "hilbert":
        return bench_hilbert(args)
    if args.workload == "timeaxis":
        return bench_timeaxis(args)

    from cuda_emd import synth

    x = synth.seismic_traces(args.batch, args.n, seed=1)
    B, N = x.shape
    # CPU baseline first: its worker processes fork before this process touches the GPU
    cpu, cpu_records = (None, None) if args.no_cpu else cpu_baseline(x, args.cpu_sample)

    import torch
    from cuda_emd import EMDKernelConfig

    dx = torch.from_numpy(x).cuda()
    offsets = torch.arange(0, (B + 1) * N, N, dtype=torch.int64)
    lengths = np.full(B, N, dtype=np.int64)
    cfg = EMDKernelConfig.for_device()
    launcher = cfg.build()
    results = {}

    events = []

    def sift_call(call):
        # HIP events on the launch stream (torch's current stream, which the launcher uses)
        # around the sift launch; read after the timed loop, so no step waits on them
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        out = call()
        e1.record()
        events.append((e0, e1))
        return out

    def timed(fn, pipe=None):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        events.clear()
        if pipe is not None:
            pipe.launch_ms()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            out = fn()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        kernel_ms = (pipe.launch_ms() if pipe is not None
                     else [a.elapsed_time(b) for a, b in events])
        return out, wall / args.steps * 1e3, float(np.mean(kernel_ms))

    if args.mode in ("both", "plain"):
        out, ms_wall, ms_ev = timed(lambda: sift_call(lambda: launcher.run(dx)))
        n_imfs = out[1].cpu().numpy()
        stats = launcher.last_stats.cpu().numpy()
        results["plain"] = dict(ms=ms_wall, kernel_ms=ms_ev, n_imfs=n_imfs, stats=stats,
                                out=out)
    if args.mode in ("both", "fifo"):
        def fifo():
            return sift_call(lambda: launcher.run_ragged(dx.view(-1), offsets,
                                                         threads=args.threads, persistent=True))
        out, ms_wall, ms_ev = timed(fifo)
        results["fifo"] = dict(ms=ms_wall, kernel_ms=ms_ev, n_imfs=out[2].cpu().numpy(),
                               stats=out[7].cpu().numpy(), out=out)
    model_info = None
    if args.mode in ("both", "packed"):
        from cuda_emd import costmodel
        from cuda_emd.packed import PipelinedEMD
        cm, model_info = train_cost_model(launcher, args)
        doff = offsets.cuda()
        plan_tile = []

        dlen = torch.from_numpy(lengths.astype(np.float64)).cuda()

        plan_stream = None if args.serial_plan else torch.cuda.Stream()

        def packed():
            # the whole learned path: features -> MLP -> queue order + tile -> sift launch.
            # The plan runs on its own stream: it reads only the input, so it need not wait
            # for the previous step's sift, and its one host read (the tile size) waits for
            # the plan alone.  Its kernels take CU slots as the previous sift's workgroups
            # retire.
            main = torch.cuda.current_stream()
            if plan_stream is None:
                f = costmodel.features(dx.view(-1), doff)
                plan = cm.plan(f, dlen, nbsym=cfg.nbsym)
            else:
                with torch.cuda.stream(plan_stream):
                    f = costmodel.features(dx.view(-1), doff)
                    plan = cm.plan(f, dlen, nbsym=cfg.nbsym)
                main.wait_stream(plan_stream)
                plan.order.record_stream(main)
            tile = args.tile or plan.tile_slots
            plan_tile.append(tile)
            return sift_call(lambda: launcher.run_ragged(dx.view(-1), offsets, order=plan.order,
                                                         threads=args.threads, tile_slots=tile,
                                                         persistent=True))

        pipe = None
        if not args.serial_steps:
            # Consecutive steps overlap: each step's sift runs on the next of two lanes
            # (handle + stream) and takes the CU slots the previous step's sift frees as its
            # queue drains.  Every step still plans and decomposes its whole batch.
            pipe = PipelinedEMD(cfg, cm, threads=args.threads, tile_slots=args.tile,
                                depth=args.lanes, timing=True)

            def packed():  # noqa: F811
                out = pipe.submit(dx.view(-1), offsets, offsets_dev=doff, lengths_dev=dlen)
                plan_tile.append(pipe.last_tile)
                return out
        out, ms_wall, ms_ev = timed(packed, pipe)
        results["packed"] = dict(ms=ms_wall, kernel_ms=ms_ev, n_imfs=out[2].cpu().numpy(),
                                 stats=out[7].cpu().numpy(), out=out)
        if pipe is not None:
            # Overlapped launches wait for CU slots the previous step still holds, so their
            # events span more than the kernel's own run.  The kernel's own duration comes
            # from the same launch (the plan's order and tile) run alone, events on its
            # stream, after the timed loop.
            order = cm.order(costmodel.features(dx.view(-1), doff) if cm.orders_by_sifts
                             else None, dlen)
            solo = []
            for _ in range(3):
                torch.cuda.synchronize()
                sift_call(lambda: launcher.run_ragged(dx.view(-1), offsets, order=order,
                                                      threads=args.threads,
                                                      tile_slots=pipe.last_tile,
                                                      persistent=True))
                torch.cuda.synchronize()
                a, b = events[-1]
                solo.append(a.elapsed_time(b))
            results["packed"]["launch_span_ms"] = ms_ev
            results["packed"]["kernel_ms"] = float(np.mean(solo))
        model_info["tile_slots"] = plan_tile[-1]
        true = np.stack([results["packed"]["stats"][:, 0], results["packed"]["n_imfs"],
                         results["packed"]["stats"][:, 1]], axis=1).astype(np.float64)
        model_info["mape_on_workload"] = costmodel.mape(
            cm.predict(costmodel.features(dx.view(-1), doff)), true)

    if args.mode == "both":
        from cuda_emd.packed import Autotuner, PipelinedEMD
        # Same-shape comparators.  plain_same_shape: the packed path's launch shape (its
        # workgroup size and LDS tile, so the same workgroups per CU) with one workgroup
        # per trace instead of the persistent queue.  packed_no_model: the same pipelined
        # persistent path in batch order with the tile the autotuner's "largest tile for 4
        # workgroups per CU" rule gives, i.e. everything but the cost model.
        tile_used = int(plan_tile[-1]) if plan_tile else 0
        out, ms_wall, ms_ev = timed(lambda: sift_call(lambda: launcher.run_ragged(
            dx.view(-1), offsets, threads=args.threads, tile_slots=tile_used,
            persistent=False)))
        results["plain_same_shape"] = dict(ms=ms_wall, kernel_ms=ms_ev,
                                           n_imfs=out[2].cpu().numpy(),
                                           stats=out[7].cpu().numpy(), out=out, tile=tile_used)
        tile4 = Autotuner._largest_tile_for(launcher, args.threads, N, 4)
        pipe0 = PipelinedEMD(cfg, None, threads=args.threads, tile_slots=tile4,
                             depth=args.lanes, timing=True)
        out, ms_wall, ms_ev = timed(lambda: pipe0.submit(dx.view(-1), offsets, offsets_dev=doff,
                                                         lengths_dev=dlen), pipe0)
        results["packed_no_model"] = dict(ms=ms_wall, kernel_ms=ms_ev,
                                          n_imfs=out[2].cpu().numpy(),
                                          stats=out[7].cpu().numpy(), out=out, tile=tile4)

    head = next(k for k in ("packed", "fifo", "plain") if k in results)
    r = results[head]
    flops = algorithmic_flops(r["stats"], r["n_imfs"], lengths)
    achieved = flops / (r["kernel_ms"] * 1e-3) / 1e12
    # overlapped steps keep the device busy with sift work back to back: the sustained rate
    sustained = flops / (r["ms"] * 1e-3) / 1e12
    line = {
        "metric": METRIC,
        "value": B / (r["ms"] * 1e-3),
        "unit": "traces/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "seismic traces, 16384 x N=2001, seed 1", "traces": B, "N": N,
                   "path": head, "max_imfs": cfg.max_imfs, "max_extrema": cfg.max_extrema,
                   "steps_overlap": (head == "packed" and not args.serial_steps)},
        # the compute roof: MI355X's fp64 peak (78.6 TFLOP/s, the same for its vector ALUs,
        # which run the sift, as for its matrix cores, which it does not use)
        "roofline": {"bound": "valu", "roof": "fp64 VALU (78.6 TFLOP/s; the sift issues no MFMA)", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic() if head == "packed" and B == 16384 else None,
                     "flops_per_launch": flops,
                     "kernel_ms": r["kernel_ms"],
                     "achieved_sustained": sustained,
                     "algorithmic_hbm_bytes": algorithmic_bytes(r["n_imfs"], lengths)},
    }
    if head == "packed" and B == 16384:
        # Hardware counters of the same launch (committed rocprofv3 passes): counted fp64
        # operations over this run's kernel time, LDS activity, and resident-slot occupancy.
        pv = profile_values()
        if "counted_fp64_ops_per_launch" in pv:
            kernel_ms = r["kernel_ms"]
            counted = pv["counted_fp64_ops_per_launch"] / (kernel_ms * 1e-3) / 1e12
            line["counters"] = {
                "counted_fp64_tflops": counted,
                "counted_fp64_frac": counted / FP64_PEAK_TFLOPS,
                "fp64_issue_per_cu_cycle": pv.get("fp64_issue_per_cu_cycle"),
                "valu_issue_per_cu_cycle": pv.get("valu_issue_per_cu_cycle"),
                "lds_active_frac": pv.get("lds_active_frac"),
                "lds_bank_conflict_frac": pv.get("lds_bank_conflict_frac"),
                "lds_tb_per_s": pv.get("lds_bytes_per_launch", 0.0) / (kernel_ms * 1e-3) / 1e12,
                "hbm_tb_per_s": pv.get("traffic_bytes_per_launch", 0.0) / (kernel_ms * 1e-3) / 1e12,
                "hbm_frac": pv.get("traffic_bytes_per_launch", 0.0) / (kernel_ms * 1e-3) / 1e9
                / HBM_PEAK_GBS,
                "mean_waves_per_cu": pv.get("mean_waves_per_cu"),
                "resident_waves_per_cu": pv.get("resident_waves_per_cu"),
                "queue_idle_frac": pv.get("queue_idle_frac"),
                "source": "profiles/emd_sift_kernels_seismic_16384x2001.md",
            }
    if "plain" in results:
        line["plain_traces_per_s"] = B / (results["plain"]["ms"] * 1e-3)
        line["plain_ms_per_step"] = results["plain"]["ms"]
    if "fifo" in results:
        line["fifo_traces_per_s"] = B / (results["fifo"]["ms"] * 1e-3)
    if "packed" in results and "plain" in results:
        line["speedup_vs_plain"] = results["plain"]["ms"] / results["packed"]["ms"]
        a, b = results["plain"]["out"], results["packed"]["out"]
        line["identical_to_plain"] = bool(
            np.array_equal(a[1].cpu().numpy(), b[2].cpu().numpy())
            and np.array_equal(a[3].cpu().numpy(), b[4].cpu().numpy())
            and np.array_equal(a[4].cpu().numpy(), b[5].cpu().numpy())
            and np.array_equal(a[5].cpu().numpy(), b[6].cpu().numpy())
            and np.array_equal(results["plain"]["stats"][:, :3], r["stats"][:, :3]))
    if "packed" in results and "fifo" in results:
        line["speedup_vs_fifo"] = results["fifo"]["ms"] / results["packed"]["ms"]
    for name in ("plain_same_shape", "packed_no_model"):
        if name in results and "packed" in results:
            q = results[name]
            line[name] = {"traces_per_s": B / (q["ms"] * 1e-3), "ms_per_step": q["ms"],
                          "kernel_ms": q["kernel_ms"], "tile_slots": q["tile"],
                          "threads": args.threads,
                          "identical": bool(all(np.array_equal(
                              q["out"][k].cpu().numpy(), results["packed"]["out"][k].cpu().numpy())
                              for k in range(2, 7)) and np.array_equal(
                              q["stats"][:, :3], results["packed"]["stats"][:, :3]))}
    if "plain_same_shape" in results and "packed" in results:
        line["speedup_vs_plain_same_shape"] = (results["plain_same_shape"]["ms"]
                                               / results["packed"]["ms"])
    if "packed_no_model" in results and "packed" in results:
        line["speedup_vs_no_model"] = results["packed_no_model"]["ms"] / results["packed"]["ms"]
    if cpu_records is not None and "packed" in results:
        line["parity_vs_oracle"] = parity_vs_oracle(x, results["packed"]["out"], lengths,
                                                    cpu_records)
    if model_info is not None:
        line["cost_model"] = model_info
    line["sifts_per_trace"] = float(r["stats"][:, 0].mean())
    line["imfs_per_trace"] = float(r["n_imfs"].mean())
    if cpu is not None:
        line["cpu_baseline"] = cpu
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
