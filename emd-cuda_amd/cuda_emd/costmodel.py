"""costmodel.py -- the learned cost model that packs the persistent queue.

A 12 -> 64 -> 64 -> 3 FP32 MLP (exact-erf GELU) predicts, per trace, the total sift
iterations, the IMF count and the peak extrema count from cheap per-trace features.  Every
step runs in the gfx950 library's own kernels: ``emd_features`` (features),
``emd_mlp_forward`` (inference), ``emd_mlp_backward`` (fused forward + backward with a
fixed-order gradient reduction) and ``emd_mlp_adam``.  It is trained on the counts the
decomposition itself records (stats: sifts, peak extrema; n_imfs), and used only to

* order the queue longest-predicted-first (LPT on predicted sifts x N), and
* size the LDS knot tile (a high quantile of the predicted peak extrema),

so decompositions are identical with or without it; FIFO order and the capacity-rule tile
are the fallback.  The order uses the predicted sifts only when they rank a held-out part
of the training traces (every eighth) with a rank correlation of at least RANK_MIN;
otherwise it orders by length alone (a stable sort: batch order among equal lengths).
On synthetic seismic traces of one length the sift count is not predictable from these
features (held-out rank correlation near 0), and an order by such predictions runs the
sift launch slower than batch order.  No reference counterpart: the reference launches one block per signal
in batch order (kernels.py:1214-1220) and sizes LDS from max_extrema (kernels.py:1205-1271).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib

N_FEATURES = 12
N_PARAMS = 5187
TARGETS = ("sifts", "n_imfs", "peak_extrema")
RANK_MIN = 0.3  # held-out rank correlation the sift prediction needs to order the queue


def _ranks(x) -> np.ndarray:
    """Ranks 0..n-1 with ties given their average rank."""
    x = np.asarray(x, dtype=np.float64)
    order = np.argsort(x, kind="mergesort")
    _, first, counts = np.unique(x[order], return_index=True, return_counts=True)
    r = np.empty(len(x))
    r[order] = np.repeat(first + (counts - 1) / 2.0, counts)
    return r


def rank_corr(a, b) -> float:
    """Spearman rank correlation (nan when either side is constant or empty)."""
    if len(a) < 2:
        return float("nan")
    ra, rb = _ranks(a), _ranks(b)
    ra, rb = ra - ra.mean(), rb - rb.mean()
    den = np.sqrt(np.sum(ra * ra) * np.sum(rb * rb))
    return float(np.sum(ra * rb) / den) if den > 0 else float("nan")


def _torch():
    import torch
    return torch


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr() if t is not None and t.numel() > 0 else 0)


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def features(values, offsets):
    """[B,12] float32 features of CSR traces (device float64 values, int64 offsets)."""
    torch = _torch()
    v = values if isinstance(values, torch.Tensor) else torch.from_numpy(np.asarray(values))
    off = offsets if isinstance(offsets, torch.Tensor) else torch.from_numpy(
        np.asarray(offsets, dtype=np.int64))
    if v.dtype != torch.float64 or off.dtype != torch.int64:
        raise ValueError("features need float64 values and int64 offsets")
    v, off = v.cuda().contiguous(), off.cuda().contiguous()
    B = off.numel() - 1
    out = torch.empty((max(B, 0), N_FEATURES), dtype=torch.float32, device=v.device)
    _lib.check(_lib.load().emd_features(_ptr(v), _ptr(off), B, _ptr(out), _stream()),
               "emd_features")
    return out


def targets_from(stats, n_imfs):
    """Raw targets [B,3] (sifts, n_imfs, peak extrema) from a decomposition's records."""
    s = stats.cpu().numpy() if hasattr(stats, "cpu") else np.asarray(stats)
    n = n_imfs.cpu().numpy() if hasattr(n_imfs, "cpu") else np.asarray(n_imfs)
    return np.stack([s[:, 0], n, s[:, 1]], axis=1).astype(np.float64)


def _transform(raw):
    return np.stack([np.log1p(raw[:, 0]), raw[:, 1], np.log1p(raw[:, 2])], axis=1)


def mape(pred, true) -> dict:
    """Mean absolute percentage error per target over traces whose true value is > 0."""
    out = {}
    for k, name in enumerate(TARGETS):
        m = true[:, k] > 0
        out[name] = float(np.mean(np.abs(pred[m, k] - true[m, k]) / true[m, k])) if m.any() \
            else float("nan")
    return out


@dataclass
class Plan:
    order: object            # int32 device tensor, queue order (LPT)
    tile_slots: int          # LDS knot slots


class CostModel:
    """Parameters, Adam moments and input/target standardisation live on the device."""

    def __init__(self, seed: int = 0, device=None):
        torch = _torch()
        self.device = torch.device(device or "cuda")
        rng = np.random.default_rng(seed)
        parts = []
        for fan_out, fan_in in ((64, 12), (64, 64), (3, 64)):
            bound = 1.0 / np.sqrt(fan_in)
            parts.append(rng.uniform(-bound, bound, fan_out * fan_in))
            parts.append(rng.uniform(-bound, bound, fan_out))
        self.params = torch.tensor(np.concatenate(parts), dtype=torch.float32, device=self.device)
        assert self.params.numel() == N_PARAMS == _lib.load().emd_mlp_param_count()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step = 0
        self.norm = torch.cat([torch.zeros(12), torch.ones(12)]).float().to(self.device)
        self.tmu = np.zeros(3)
        self.tsd = np.ones(3)
        self.history: list[float] = []
        self.sift_rank = float("nan")  # held-out rank correlation of predicted sifts
        self._tstd = None  # (tsd, tmu) on the device, made by the first plan after a fit/load

    @property
    def orders_by_sifts(self) -> bool:
        """Whether the predicted sifts enter the queue order (no held-out evidence: yes)."""
        return not (self.sift_rank < RANK_MIN)

    # -- kernels -------------------------------------------------------------------------
    def forward_std(self, feats):
        torch = _torch()
        f = feats.to(self.device, torch.float32).contiguous()
        out = torch.empty((f.shape[0], 3), dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().emd_mlp_forward(_ptr(self.params), _ptr(self.norm), _ptr(f),
                                                f.shape[0], _ptr(out), _stream()),
                   "emd_mlp_forward")
        return out

    def loss_grad(self, feats, targets_std, target_weight):
        """(grad [N_PARAMS+1] with the loss last) of one mini-batch."""
        torch = _torch()
        M = feats.shape[0]
        ws = torch.empty((int(_lib.load().emd_mlp_workspace_floats(M)),), dtype=torch.float32,
                         device=self.device)
        grad = torch.empty((N_PARAMS + 1,), dtype=torch.float32, device=self.device)
        _lib.check(_lib.load().emd_mlp_backward(
            _ptr(self.params), _ptr(self.norm), _ptr(feats), _ptr(targets_std),
            _ptr(target_weight), M, _ptr(ws), _ptr(grad), _stream()), "emd_mlp_backward")
        return grad

    def adam(self, grad, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        _lib.check(_lib.load().emd_mlp_adam(_ptr(self.params), _ptr(self.m), _ptr(self.v),
                                             _ptr(grad), self.step, lr, beta1, beta2, eps,
                                             _stream()), "emd_mlp_adam")

    # -- training / inference ---------------------------------------------------------------
    def fit(self, feats, raw_targets, epochs: int = 40, batch: int = 2048, lr: float = 3e-3,
            seed: int = 0, target_weight=(1.0, 1.0, 1.0)):
        """Mini-batch Adam on the standardised transformed targets (log1p sifts, n_imfs,
        log1p peak), on all traces but every eighth (when there are 64 or more), which is
        held out to measure how well the predicted sifts rank (`sift_rank`).  Returns the
        per-epoch mean loss."""
        torch = _torch()
        f_all = feats.to(self.device, torch.float32).contiguous()
        raw_all = np.asarray(raw_targets, dtype=np.float64)
        hold = (np.arange(f_all.shape[0]) % 8 == 7) if f_all.shape[0] >= 64 else \
            np.zeros(f_all.shape[0], dtype=bool)
        keep = torch.from_numpy(np.nonzero(~hold)[0]).to(self.device)
        f = f_all[keep].contiguous()
        raw_targets = raw_all[~hold]
        M = f.shape[0]
        mu = f.mean(0)
        sd = f.std(0).clamp_min(1e-3)
        self.norm = torch.cat([mu, sd]).contiguous()
        t = _transform(np.asarray(raw_targets, dtype=np.float64))
        self.tmu, self.tsd = t.mean(0), np.maximum(t.std(0), 1e-3)
        self._tstd = None
        ts = torch.tensor((t - self.tmu) / self.tsd, dtype=torch.float32, device=self.device)
        tw = torch.tensor(target_weight, dtype=torch.float32, device=self.device)
        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        for _ in range(epochs):
            perm = torch.randperm(M, device=self.device, generator=g)
            losses = []
            for s in range(0, M, batch):
                idx = perm[s:s + batch]
                grad = self.loss_grad(f[idx].contiguous(), ts[idx].contiguous(), tw)
                self.adam(grad, lr)
                losses.append(grad[N_PARAMS:])
            self.history.append(float(torch.cat(losses).mean()))
        if hold.any():
            held = torch.from_numpy(np.nonzero(hold)[0]).to(self.device)
            pred = self.predict(f_all[held].contiguous())
            self.sift_rank = rank_corr(pred[:, 0], raw_all[hold, 0])
        return self.history

    def predict(self, feats) -> np.ndarray:
        """Raw predictions [B,3]: sifts, n_imfs, peak extrema."""
        y = self.forward_std(feats).double().cpu().numpy() * self.tsd + self.tmu
        return np.stack([np.expm1(y[:, 0]), y[:, 1], np.expm1(y[:, 2])], axis=1)

    def plan(self, feats, lengths, nbsym: int = 2, quantile: float = 0.995,
             margin: float = 1.1, cap: int | None = None) -> Plan:
        """Queue order (longest predicted sifts x N first; N alone when the sift
        prediction did not rank the held-out traces, see `orders_by_sifts`) and LDS knot
        tile (both envelopes of a sift at the ``quantile`` of predicted peak extrema,
        x margin).  Everything stays on the device except the one tile size the launch
        needs."""
        torch = _torch()
        y = self.forward_std(feats).double()
        if self._tstd is None:
            self._tstd = (torch.tensor(self.tsd, dtype=torch.float64, device=self.device),
                          torch.tensor(self.tmu, dtype=torch.float64, device=self.device))
        tsd, tmu = self._tstd
        y = y * tsd + tmu
        L = lengths if isinstance(lengths, torch.Tensor) else torch.from_numpy(
            np.asarray(lengths, dtype=np.float64))
        L = L.to(self.device, torch.float64)
        cost = (torch.expm1(y[:, 0]).clamp_min(1.0) * L) if self.orders_by_sifts else L
        order = torch.argsort(cost, descending=True, stable=True).to(torch.int32)
        if y.shape[0]:
            peak = float(torch.quantile(torch.expm1(y[:, 2]).clamp_min(0.0), quantile))
        else:
            peak = 0.0
        tile = int(2 * np.ceil(peak * margin) + 4 * nbsym)
        tile = max(tile, 16 + 4 * nbsym)
        if cap is not None:
            tile = min(tile, cap)
        return Plan(order, tile)

    def order(self, feats, lengths):
        """The queue order of plan() alone, on the device (no host read): longest predicted
        sifts x N first, or N alone when the sift prediction does not rank (then feats may
        be None)."""
        torch = _torch()
        L = lengths if isinstance(lengths, torch.Tensor) else torch.from_numpy(
            np.asarray(lengths, dtype=np.float64))
        L = L.to(self.device, torch.float64)
        if not self.orders_by_sifts:
            return torch.argsort(L, descending=True, stable=True).to(torch.int32)
        y = self.forward_std(feats).double()
        if self._tstd is None:
            self._tstd = (torch.tensor(self.tsd, dtype=torch.float64, device=self.device),
                          torch.tensor(self.tmu, dtype=torch.float64, device=self.device))
        tsd, tmu = self._tstd
        cost = torch.expm1(y[:, 0] * tsd[0] + tmu[0]).clamp_min(1.0) * L
        return torch.argsort(cost, descending=True, stable=True).to(torch.int32)

    # -- persistence -------------------------------------------------------------------------
    def state(self) -> dict:
        return {"params": self.params.cpu().numpy(), "norm": self.norm.cpu().numpy(),
                "tmu": self.tmu, "tsd": self.tsd, "sift_rank": np.array(self.sift_rank)}

    def save(self, path: str) -> None:
        np.savez(path, **self.state())

    @classmethod
    def load(cls, path: str, device=None) -> "CostModel":
        torch = _torch()
        d = np.load(path, allow_pickle=False)
        cm = cls(device=device)
        cm.params = torch.tensor(d["params"], dtype=torch.float32, device=cm.device)
        cm.norm = torch.tensor(d["norm"], dtype=torch.float32, device=cm.device)
        cm.tmu, cm.tsd = np.array(d["tmu"]), np.array(d["tsd"])
        cm._tstd = None
        cm.sift_rank = float(d["sift_rank"]) if "sift_rank" in d.files else float("nan")
        return cm
