"""CPU restatement (fp64 NumPy) of the cost model's per-trace features and its MLP
(12 -> 64 -> 64 -> 3, exact-erf GELU): forward, loss gradient and one Adam step.

TEST INFRASTRUCTURE ONLY: tests/ use it as the checker of the HIP kernels in
emd-cuda_amd/csrc/emd_costmodel.hip; the product path never imports it.

The cost model has no counterpart in the reference (its launch is one block per signal,
kernels.py:1214-1220), so nothing here follows a reference file.  Pinning (tests/
test_costmodel_oracle.py): features against a direct-definition DFT and brute-force
counting loops and a pure tone at a known bin; forward/backward against torch's own
Linear/GELU/autograd in float64 and against central finite differences; Adam against
torch.optim.Adam.

Parameter layout (flat): W1[64][12] | b1[64] | W2[64][64] | b2[64] | W3[3][64] | b3[3].
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

N_FEATURES = 12
HIDDEN = 64
N_OUT = 3
SEG = 256
BAND_EDGES = (1, 2, 4, 8, 16, 32, 64, 96, 129)
SHAPES = (("W1", (HIDDEN, N_FEATURES)), ("b1", (HIDDEN,)), ("W2", (HIDDEN, HIDDEN)),
          ("b2", (HIDDEN,)), ("W3", (N_OUT, HIDDEN)), ("b3", (N_OUT,)))
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)  # 5187


# ---------------------------------------------------------------------------------------
# features
# ---------------------------------------------------------------------------------------
def band_power(x: np.ndarray) -> np.ndarray:
    """Mean |rfft|^2 (129 bins) over non-overlapping 256-sample segments; one zero-padded
    segment when len(x) < 256 (the tail shorter than a segment is not used)."""
    n = len(x)
    if n >= SEG:
        k = n // SEG
        segs = x[:k * SEG].reshape(k, SEG)
    else:
        segs = np.zeros((1, SEG))
        segs[0, :n] = x
    return np.mean(np.abs(np.fft.rfft(segs, axis=1)) ** 2, axis=0)


def features(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    f = np.zeros(N_FEATURES)
    f[0] = np.log(max(n, 1))
    d = np.diff(x)
    f[1] = np.log1p(np.count_nonzero(d[:-1] * d[1:] < 0)) if n >= 3 else 0.0
    f[2] = np.log1p(np.count_nonzero(x[:-1] * x[1:] < 0)) if n >= 2 else 0.0
    f[3] = np.log(np.sqrt(np.mean(x * x)) + 1e-12) if n else np.log(1e-12)
    p = band_power(x)
    tot = p.sum()
    for b in range(8):
        e = p[BAND_EDGES[b]:BAND_EDGES[b + 1]].sum()
        f[4 + b] = np.log((e / tot if tot > 0 else 0.0) + 1e-10)
    return f


def features_batch(values: np.ndarray, offsets: np.ndarray) -> np.ndarray:
    B = len(offsets) - 1
    return np.stack([features(values[offsets[b]:offsets[b + 1]]) for b in range(B)]) if B \
        else np.zeros((0, N_FEATURES))


# ---------------------------------------------------------------------------------------
# MLP
# ---------------------------------------------------------------------------------------
def unpack(params: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        k = int(np.prod(shape))
        out[name] = params[o:o + k].reshape(shape)
        o += k
    return out


def pack(parts: dict) -> np.ndarray:
    return np.concatenate([np.asarray(parts[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def forward(params, norm, F):
    """y[M,3] and the activations backward needs."""
    P = unpack(np.asarray(params, dtype=np.float64))
    norm = np.asarray(norm, dtype=np.float64)
    X = (np.asarray(F, dtype=np.float64) - norm[:N_FEATURES]) / norm[N_FEATURES:]
    z1 = X @ P["W1"].T + P["b1"]
    h1 = gelu(z1)
    z2 = h1 @ P["W2"].T + P["b2"]
    h2 = gelu(z2)
    y = h2 @ P["W3"].T + P["b3"]
    return y, (X, z1, h1, z2, h2)


def loss_and_grad(params, norm, F, T, tw):
    """L = (1/M) sum_s sum_o tw[o] (y_so - T_so)^2 and dL/dparams (flat)."""
    y, (X, z1, h1, z2, h2) = forward(params, norm, F)
    P = unpack(np.asarray(params, dtype=np.float64))
    M = y.shape[0]
    tw = np.asarray(tw, dtype=np.float64)
    e = y - np.asarray(T, dtype=np.float64)
    L = float(np.sum(tw * e * e) / M)
    dy = 2.0 * tw * e / M
    g = {"W3": dy.T @ h2, "b3": dy.sum(0)}
    dz2 = (dy @ P["W3"]) * gelu_grad(z2)
    g["W2"], g["b2"] = dz2.T @ h1, dz2.sum(0)
    dz1 = (dz2 @ P["W2"]) * gelu_grad(z1)
    g["W1"], g["b1"] = dz1.T @ X, dz1.sum(0)
    return L, pack(g)


def adam(params, m, v, grad, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam step; returns new (params, m, v)."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
