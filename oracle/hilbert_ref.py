"""CPU restatement (NumPy/SciPy) of the Hilbert spectral step: instantaneous amplitude,
phase and frequency of an IMF row from its analytic signal.

TEST INFRASTRUCTURE ONLY: tests/ and bench.py use it as the checker of the HIP kernel in
emd-cuda_amd/csrc/emd_hilbert.hip; the product path never imports it.

The reference has no Hilbert step (it stops at the IMFs, kernels.py:1343-1372, of traces
it describes as seismic, sampled at 4 ms, emd.py:28-30), so nothing here follows a
reference file.  The analytic signal is ``scipy.signal.hilbert`` (scipy as installed in
this image), and the attributes follow the recipe of scipy's own hilbert documentation:
amplitude ``abs(z)``, phase ``angle(z)``, frequency ``diff(unwrap(angle(z))) / (2 pi dt)``;
the last sample repeats the one before so the frequency row keeps the row's length.

``hilbert_conv`` restates the formulation the HIP kernel computes for a row of any length
N: the circular convolution of the row with the closed-form N-point discrete Hilbert
kernel, taken through a zero-padded power-of-two FFT of at least 2N-1 points.  It is not
the checker; the tests pin it (and the kernel's host emulation) to scipy.

Pinning (tests/test_hilbert_oracle.py): bin-centred tones, whose analytic signal is an
exact complex exponential (amplitude, phase and constant frequency known in closed form);
bin-centred amplitude-modulated tones (Bedrosian: the envelope comes back exactly); the
N-point kernel's impulse response against its definition as the inverse DFT of
-i sgn(k); ``hilbert_conv`` against scipy for every N up to 300 and long odd/even/prime
lengths.
"""
from __future__ import annotations

import numpy as np
import scipy.signal


def analytic(x: np.ndarray, n_fft: int = 0) -> np.ndarray:
    """z = x + i H x of a real row: ``scipy.signal.hilbert(x)`` for n_fft == 0, else
    ``scipy.signal.hilbert(x, N=L)[:len(x)]`` with L = n_fft, or for n_fft == -1 the
    smallest power of two >= len(x)."""
    x = np.asarray(x, dtype=np.float64)
    n = x.shape[-1]
    if n == 0:
        return x.astype(np.complex128)
    if n_fft == 0:
        return scipy.signal.hilbert(x, axis=-1)
    L = (1 << int(np.ceil(np.log2(n)))) if n_fft < 0 else int(n_fft)
    return scipy.signal.hilbert(x, N=L, axis=-1)[..., :n]


def instantaneous(x: np.ndarray, dt: float = 1.0, n_fft: int = 0):
    """(amp, phase, freq) of a row (or rows along the last axis), as float64 arrays of the
    row's shape."""
    z = analytic(x, n_fft)
    amp = np.abs(z)
    phase = np.angle(z)
    n = z.shape[-1]
    if n <= 1:
        return amp, phase, np.zeros_like(amp)
    f = np.diff(np.unwrap(phase, axis=-1), axis=-1) / (2.0 * np.pi * dt)
    freq = np.concatenate([f, f[..., -1:]], axis=-1)
    return amp, phase, freq


def kernel_value(n: np.ndarray, N: int) -> np.ndarray:
    """Impulse response of the N-point multiplier -i sgn(k) at 1 <= n <= N-1 (half-angle
    closed forms): N odd: cot(pi n / 2N) / N for odd n, -tan(pi n / 2N) / N for even n;
    N even: 2 cot(pi n / N) / N for odd n, 0 for even n."""
    n = np.asarray(n, dtype=np.int64)
    odd = (n % 2) == 1
    if N % 2:
        th = np.pi * n / (2.0 * N)
        return np.where(odd, 1.0 / np.tan(th), -np.tan(th)) / N
    with np.errstate(divide="ignore"):
        return np.where(odd, (2.0 / N) / np.tan(np.pi * n / N), 0.0)


def hilbert_conv(x: np.ndarray) -> np.ndarray:
    """H x of one row by the kernel's formulation: circular convolution with the N-point
    kernel, laid out on M = 2^ceil(log2(2N-1)) points so that a length-M circular
    convolution reproduces it on the first N outputs."""
    x = np.asarray(x, dtype=np.float64)
    N = len(x)
    if N <= 2:
        return np.zeros(N)
    M = 1 << int(np.ceil(np.log2(2 * N - 1)))
    g = np.zeros(M)
    j = np.arange(1, N)
    g[j] = kernel_value(j, N)
    g[M - j] = -kernel_value(j, N)
    xp = np.zeros(M)
    xp[:N] = x
    return np.real(np.fft.ifft(np.fft.fft(xp) * np.fft.fft(g)))[:N]
