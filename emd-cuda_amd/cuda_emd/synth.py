"""Seeded synthetic trace generators for the five benchmark workloads.

Every generator uses ``np.random.default_rng(seed)`` and draws in the order written in
its docstring, so the same seed always yields the same traces.  Ragged workloads are
returned in CSR form: ``values`` (float64, all traces back to back) and ``offsets``
(int64, length B+1).
"""
from __future__ import annotations

import numpy as np
from scipy import signal as sps

DT = 0.004          # 4 ms sampling of the seismic traces
FS = 1.0 / DT       # 250 Hz


def chirp_mix(B: int = 256, N: int = 1024, seed: int = 0) -> np.ndarray:
    """Three-component test signals (workload "chirp-mix", B=256 x N=1024, seed 0).

    x[n] = a1 sin(2 pi f1 n + phi1) + a2 sin(2 pi f2 n + phi2)
           + a3 sin(2 pi (f3 n + 0.5 k n^2)) + 0.05 N(0,1)
    Draw order: a [B,3] ~ U(0.5,2); f1, f2, f3 [B] ~ U(.002,.01), U(.02,.05), U(.08,.15);
    k [B] ~ U(0, 5e-5); phi [B,2] ~ U(0, 2 pi); noise [B,N] ~ N(0,1).
    """
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.5, 2.0, (B, 3))
    f1 = rng.uniform(0.002, 0.01, B)
    f2 = rng.uniform(0.02, 0.05, B)
    f3 = rng.uniform(0.08, 0.15, B)
    k = rng.uniform(0.0, 5e-5, B)
    phi = rng.uniform(0.0, 2 * np.pi, (B, 2))
    noise = rng.standard_normal((B, N))
    n = np.arange(N, dtype=np.float64)[None, :]
    x = (a[:, :1] * np.sin(2 * np.pi * f1[:, None] * n + phi[:, :1])
         + a[:, 1:2] * np.sin(2 * np.pi * f2[:, None] * n + phi[:, 1:2])
         + a[:, 2:3] * np.sin(2 * np.pi * (f3[:, None] * n + 0.5 * k[:, None] * n * n))
         + 0.05 * noise)
    return np.ascontiguousarray(x)


def ricker(f: float = 25.0, half_width_s: float = 0.1, dt: float = DT) -> np.ndarray:
    """Ricker wavelet (1 - 2 pi^2 f^2 t^2) exp(-pi^2 f^2 t^2) on t in [-0.1 s, 0.1 s]."""
    t = np.arange(-half_width_s, half_width_s + dt / 2, dt)
    a = (np.pi * f * t) ** 2
    return (1.0 - 2.0 * a) * np.exp(-a)


_BP_SOS = None


def _bandpass_sos():
    global _BP_SOS
    if _BP_SOS is None:
        _BP_SOS = sps.butter(4, [5.0, 60.0], btype="bandpass", fs=FS, output="sos")
    return _BP_SOS


def seismic_traces(B: int = 16384, N: int = 2001, seed: int = 1, rng=None) -> np.ndarray:
    """Synthetic seismic traces (workload "seismic", B=16384 x N=2001, seed 1).

    Sparse reflectivity (Bernoulli p=0.02 spikes with N(0,1) amplitudes) convolved
    ('same' length) with a 25 Hz Ricker wavelet sampled at 4 ms over +-0.1 s, plus N(0,1)
    noise band-passed 5-60 Hz (4th-order Butterworth, zero phase) scaled per trace to
    10 dB SNR (signal power / noise power = 10).
    Draw order: spikes [B,N] ~ U(0,1) < 0.02; amplitudes [B,N] ~ N(0,1); noise [B,N].
    """
    rng = rng if rng is not None else np.random.default_rng(seed)
    spikes = rng.uniform(0.0, 1.0, (B, N)) < 0.02
    amps = rng.standard_normal((B, N))
    refl = np.where(spikes, amps, 0.0)
    w = ricker()
    clean = sps.fftconvolve(refl, w[None, :], mode="same", axes=1)
    noise = sps.sosfiltfilt(_bandpass_sos(), rng.standard_normal((B, N)), axis=1)
    ps = np.mean(clean * clean, axis=1, keepdims=True)
    pn = np.mean(noise * noise, axis=1, keepdims=True)
    scale = np.sqrt(ps / (10.0 * np.maximum(pn, 1e-300)))
    return np.ascontiguousarray(clean + scale * noise)


def seismic_chunks(total: int, seed: int, N: int = 2001, chunk: int = 16384,
                   progress=None) -> np.ndarray:
    """``total`` seismic traces (as seismic_traces) drawn chunk by chunk of ``chunk``
    traces from one generator seeded ``seed`` (the cost-model workload's training and
    held-out seismic traces)."""
    rng = np.random.default_rng(seed)
    parts, left = [], total
    while left > 0:
        parts.append(seismic_traces(min(chunk, left), N, rng=rng))
        left -= len(parts[-1])
        if progress is not None:
            progress(f"seismic seed {seed}: {total - left}/{total}")
    return np.concatenate(parts) if parts else np.zeros((0, N))


def heldout_mix(total: int = 262144, seed: int = 4, white_frac: float = 0.05,
                white_N: int = 8192, progress=None):
    """Held-out set of the cost-model workload (262144 traces, seed 4): seismic traces
    (N=2001, seismic_chunks with ``seed``) plus ``white_frac`` N(0,1) white noise of length
    ``white_N``, shuffled.  A second generator seeded ``seed`` draws the white traces and
    then the permutation.  Returns (values, offsets, kinds) with kinds 0=seismic 1=white."""
    rng = np.random.default_rng(seed)
    n_white = int(round(white_frac * total))
    xs = seismic_chunks(total - n_white, seed, progress=progress)
    traces = [xs[k] for k in range(len(xs))] + [rng.standard_normal(white_N)
                                                for _ in range(n_white)]
    kinds = np.array([0] * len(xs) + [1] * n_white, np.int8)
    perm = rng.permutation(len(traces))
    traces = [traces[k] for k in perm]
    values, offsets = _csr(traces)
    return values, offsets, kinds[perm]


def ar1(N: int, phi: float, rng) -> np.ndarray:
    """Unit-variance AR(1): x[0] ~ N(0,1), x[n] = phi x[n-1] + sqrt(1-phi^2) e[n]."""
    e = rng.standard_normal(N)
    x = np.empty(N)
    x[0] = e[0]
    s = np.sqrt(1.0 - phi * phi)
    for i in range(1, N):
        x[i] = phi * x[i - 1] + s * e[i]
    return x


def ar1_batch(B: int, N: int, phi: float, rng) -> np.ndarray:
    """B independent unit-variance AR(1) traces, x[:,0] = e[:,0] (stationary start)."""
    e = rng.standard_normal((B, N))
    s = np.sqrt(1.0 - phi * phi)
    x = sps.lfilter([s], [1.0, -phi], e[:, 1:], axis=1, zi=(phi * e[:, :1]))[0]
    return np.concatenate([e[:, :1], x], axis=1)


def gated_sinusoid(N: int, rng, f: float = 0.05) -> np.ndarray:
    """0.05 cycle/sample carrier switched on and off in alternating windows whose
    lengths are U{32..256}; the first window is on with probability 1/2.  Off windows
    are exact zeros."""
    x = np.zeros(N)
    phase = rng.uniform(0.0, 2 * np.pi)
    on = rng.uniform() < 0.5
    i = 0
    n = np.arange(N)
    while i < N:
        L = int(rng.integers(32, 257))
        if on:
            x[i:i + L] = np.sin(2 * np.pi * f * n[i:i + L] + phase)
        on = not on
        i += L
    return x


def _csr(traces):
    lengths = np.array([len(t) for t in traces], dtype=np.int64)
    offsets = np.zeros(len(traces) + 1, dtype=np.int64)
    np.cumsum(lengths, out=offsets[1:])
    values = np.concatenate(traces) if traces else np.zeros(0)
    return np.ascontiguousarray(values, dtype=np.float64), offsets


def ragged_mix(B: int = 65536, seed: int = 2, n_lo: int = 256, n_hi: int = 8192,
               white_frac: float = 0.0, white_N: int = 8192):
    """Ragged mixed-content workload (workload "ragged-mix", B=65536, seed 2).

    N = round(exp(U(ln 256, ln 8192))).  Content by a seeded permutation: 50% seismic
    traces (generated at N=2001 as in seismic_traces and FFT-resampled to N), 30% AR(1)
    phi=0.95 unit variance, 20% gated 0.05 cycle/sample sinusoids.  ``white_frac`` > 0
    replaces that fraction with N(0,1) white noise of length ``white_N`` (the held-out
    set of the cost-model workload).
    Returns (values, offsets, kinds) with kinds 0=seismic 1=AR(1) 2=gated 3=white.
    """
    rng = np.random.default_rng(seed)
    Ns = np.rint(np.exp(rng.uniform(np.log(n_lo), np.log(n_hi), B))).astype(np.int64)
    n_white = int(round(white_frac * B))
    rest = B - n_white
    n_seis = int(round(0.5 * rest))
    n_ar = int(round(0.3 * rest))
    n_gate = rest - n_seis - n_ar
    kinds = np.concatenate([np.zeros(n_seis, np.int8), np.ones(n_ar, np.int8),
                            np.full(n_gate, 2, np.int8), np.full(n_white, 3, np.int8)])
    kinds = kinds[rng.permutation(B)]
    seis = seismic_traces(n_seis, 2001, rng=rng) if n_seis else np.zeros((0, 2001))
    traces = []
    si = 0
    for b in range(B):
        kd = kinds[b]
        N = int(Ns[b])
        if kd == 0:
            traces.append(sps.resample(seis[si], N) if N != 2001 else seis[si].copy())
            si += 1
        elif kd == 1:
            traces.append(ar1(N, 0.95, rng))
        elif kd == 2:
            traces.append(gated_sinusoid(N, rng))
        else:
            Ns[b] = white_N
            traces.append(rng.standard_normal(white_N))
    values, offsets = _csr(traces)
    return values, offsets, kinds


def stress_mix(B: int = 262144, N: int = 4096, seed: int = 5):
    """Stress workload (B=262144 x N=4096, seed 5): 70% chirp-mix signals (generated at
    N=4096), 25% AR(1) phi=0.99, 5% linear FM chirps sweeping 0.001 -> 0.45
    cycle/sample plus 0.01 N(0,1).  Returns (signals [B,N], kinds)."""
    rng = np.random.default_rng(seed)
    n_c = int(round(0.7 * B))
    n_ar = int(round(0.25 * B))
    n_fm = B - n_c - n_ar
    kinds = np.concatenate([np.zeros(n_c, np.int8), np.ones(n_ar, np.int8), np.full(n_fm, 2, np.int8)])
    kinds = kinds[rng.permutation(B)]
    out = np.empty((B, N))
    sub = int(rng.integers(0, 2**31 - 1))
    out[kinds == 0] = chirp_mix(n_c, N, seed=sub)
    out[kinds == 1] = ar1_batch(n_ar, N, 0.99, rng)
    n = np.arange(N)
    f0, f1 = 0.001, 0.45
    ph = 2 * np.pi * (f0 * n + 0.5 * (f1 - f0) / (N - 1) * n * n)
    out[kinds == 2] = np.sin(ph)[None, :] + 0.01 * rng.standard_normal((n_fm, N))
    return out, kinds
