"""CPU tier of the Hilbert spectral step: the oracle (oracle/hilbert_ref.py) pinned by closed
forms, the kernel's formulation (circular convolution with the closed-form discrete Hilbert
kernel) and its host emulation (tests/native/hilbert_emu.hip: the gfx950 kernel's plan,
butterflies and index maps run lane by lane on the CPU) against scipy.signal.hilbert, and
the C ABI's argument checks (no device calls)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest
import scipy.signal

from oracle import hilbert_ref as H

HERE = os.path.dirname(os.path.abspath(__file__))
NATIVE = os.path.join(HERE, "native")
EMU = os.path.join(NATIVE, "libhilbert_emu.so")


@pytest.fixture(scope="module")
def emu():
    if not os.path.exists(EMU):
        subprocess.run(["make", "-C", NATIVE], check=True)
    lib = ctypes.CDLL(EMU)
    lib.hilbert_emu.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.c_void_p]
    lib.hilbert_emu_class.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    return lib


def _emu_h(lib, x, n_fft):
    x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
    h = np.full_like(x, np.nan)
    c = lib.hilbert_emu(x.ctypes.data, x.shape[0], x.shape[1], n_fft, h.ctypes.data)
    assert c > 0
    return h


# ---- the oracle pinned by closed forms ------------------------------------------------

@pytest.mark.parametrize("N,k0", [(64, 5), (1000, 37), (2001, 250), (127, 63), (4096, 1)])
def test_tone_is_exact_exponential(N, k0):
    """A bin-centred tone A cos(2 pi k0 n / N + p) (0 < k0 < N/2) has analytic signal
    A exp(i (2 pi k0 n / N + p)): amplitude A, phase the wrapped argument, frequency
    k0 / (N dt) at every sample (the last included)."""
    n = np.arange(N)
    A, p, dt = 2.5, 0.3, 0.004
    arg = 2 * np.pi * k0 * n / N + p
    amp, phase, freq = H.instantaneous(A * np.cos(arg), dt=dt)
    assert np.max(np.abs(amp - A)) <= 1e-12 * A
    assert np.max(np.abs(np.angle(np.exp(1j * (phase - arg))))) <= 1e-12
    assert np.max(np.abs(freq - k0 / (N * dt))) <= 1e-9 * k0 / (N * dt)


def test_am_tone_envelope_is_exact():
    """(1 + 0.5 cos(2 pi m n / N)) cos(2 pi k n / N), bin-centred with k > m and k + m < N/2:
    the spectra of envelope and carrier do not overlap, so the amplitude is the envelope."""
    N, m, k = 1024, 3, 100
    n = np.arange(N)
    env = 1 + 0.5 * np.cos(2 * np.pi * m * n / N)
    amp, _, freq = H.instantaneous(env * np.cos(2 * np.pi * k * n / N))
    assert np.max(np.abs(amp - env)) <= 1e-12
    assert np.max(np.abs(freq - k / N)) <= 1e-12


def test_padded_transform_matches_scipy_n():
    x = np.random.default_rng(3).standard_normal(300)
    z = H.analytic(x, n_fft=-1)
    assert np.allclose(z, scipy.signal.hilbert(x, N=512)[:300], rtol=0, atol=1e-13)
    assert np.allclose(H.analytic(x, n_fft=1024), scipy.signal.hilbert(x, N=1024)[:300],
                       rtol=0, atol=1e-13)


def test_frequency_is_wrapped_phase_increment():
    """The frequency is the phase increment wrapped into [-pi, pi], so a row whose phase
    advances by more than pi per sample reads as the aliased (negative) frequency."""
    N = 256
    n = np.arange(N)
    x = np.cos(2 * np.pi * 100 * n / N)
    _, _, f = H.instantaneous(x)
    assert np.allclose(f, 100 / N, atol=1e-12)
    assert np.all(H.instantaneous(np.ones(1))[2] == 0)


@pytest.mark.parametrize("N", [3, 4, 5, 8, 9, 100, 101, 2001])
def test_kernel_values_are_inverse_dft_of_multiplier(N):
    """kernel_value(n) = (1/N) sum_k (-i sgn(k)) exp(2 pi i k n / N), sgn = +1 for
    1 <= k < N/2, -1 for N/2 < k < N, 0 at DC and Nyquist."""
    k = np.arange(N)
    sgn = np.where((k > 0) & (2 * k < N), 1.0, np.where(2 * k > N, -1.0, 0.0))
    q = np.fft.ifft(-1j * sgn)
    assert np.max(np.abs(q.imag)) <= 1e-15
    n = np.arange(1, N)
    assert np.max(np.abs(H.kernel_value(n, N) - q.real[1:])) <= 1e-13
    assert abs(q.real[0]) <= 1e-15


@pytest.mark.parametrize("N", list(range(1, 301)) + [1023, 1024, 1025, 2001, 4093, 8191, 8192])
def test_convolution_form_matches_scipy(N):
    x = np.random.default_rng(N).standard_normal(N)
    ref = scipy.signal.hilbert(x).imag
    assert np.max(np.abs(H.hilbert_conv(x) - ref)) <= 1e-12 * max(1.0, np.max(np.abs(x)))


# ---- the kernel's own arithmetic, emulated on the CPU ---------------------------------

def test_kernel_atan2_accuracy(emu):
    """The kernel's atan2 (one division, a 21-term polynomial, two-double pi offsets) is
    within 2 ulp of numpy's arctan2 (itself within 0.5 ulp) on random arguments of every
    quadrant and magnitude ratio, and has C's signed-zero results."""
    emu.hilbert_emu_atan2.restype = ctypes.c_double
    emu.hilbert_emu_atan2.argtypes = [ctypes.c_double, ctypes.c_double]
    rng = np.random.default_rng(8)
    y = rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000)
    x = rng.standard_normal(20000) * 10.0 ** rng.integers(-8, 8, 20000)
    got = np.array([emu.hilbert_emu_atan2(a, b) for a, b in zip(y, x)])
    ref = np.arctan2(y, x)
    ulp = np.abs(got - ref) / np.spacing(np.abs(ref))
    assert ulp.max() <= 2.0, ulp.max()
    for a, b in [(0.0, 0.0), (0.0, -0.0), (-0.0, 0.0), (-0.0, -0.0), (1.0, 0.0), (-1.0, 0.0),
                 (0.0, -1.0), (-0.0, -1.0), (1.0, 1.0), (-2.0, -2.0)]:
        g, r = emu.hilbert_emu_atan2(a, b), np.arctan2(a, b)
        assert g == r and np.signbit(g) == np.signbit(r), (a, b, g, r)


def test_kernel_twiddles_and_kernel_values(emu):
    """exp(-2 pi i q / 2^lg) with integer argument reduction, and the discrete Hilbert
    kernel's values, against numpy at long double precision."""
    out = (ctypes.c_double * 2)()
    emu.hilbert_emu_expm.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    worst = 0.0
    for lg in (2, 3, 5, 11, 14):
        M = 1 << lg
        for q in range(0, M, max(1, M // 512)):
            emu.hilbert_emu_expm(q, lg, out)
            a = 2 * np.pi * np.longdouble(q) / M
            worst = max(worst, abs(out[0] - float(np.cos(a))), abs(out[1] + float(np.sin(a))))
    assert worst <= 4e-16, worst  # about 2 ulp of 1
    emu.hilbert_emu_kernel_value.restype = ctypes.c_double
    emu.hilbert_emu_kernel_value.argtypes = [ctypes.c_int, ctypes.c_int]
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 30
    for N in (3, 4, 101, 2001, 2048, 8191):
        n = sorted(set(range(1, min(N, 40))) | set(range(max(1, N - 40), N))
                   | set(range(1, N, max(1, N // 200))))
        got = np.array([emu.hilbert_emu_kernel_value(int(k), N) for k in n])
        # (2/N) sum_{k=1}^{ceil(N/2)-1} sin(2 pi k n / N) summed in closed form at 30 digits:
        # (cos(pi n/N) - (-1)^n) / (N sin(pi n/N)) for odd N, (1 - (-1)^n) cot(pi n/N) / N
        # for even N
        x = [mp.mpf(k) / N for k in n]
        if N % 2:
            ref = [(mp.cospi(v) - (-1) ** k) / (N * mp.sinpi(v)) for k, v in zip(n, x)]
        else:
            ref = [(1 - (-1) ** k) * mp.cospi(v) / (N * mp.sinpi(v)) for k, v in zip(n, x)]
        ref = np.array([float(r) for r in ref])
        assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0 / N)) <= 1e-15, N
        assert np.max(np.abs(H.kernel_value(np.array(n), N) - ref)
                      / np.maximum(np.abs(ref), 1.0 / N)) <= 1e-12, N


def test_emulator_stage_maps(emu):
    """Every stage's (lane, register) -> position map is a bijection of [0, F) whose padded
    LDS index stays in the exchange buffer, and the digit reversal is a bijection."""
    for logf in range(2, 15):
        assert emu.hilbert_emu_check_maps(logf) == 0, logf


def test_transform_classes(emu):
    ex = ctypes.c_int(0)
    cases = {  # (len, n_fft) -> (class, exact)
        (1, 0): (1, 0), (2, 0): (1, 0), (3, 0): (3, 1), (4, 0): (2, 0), (2001, 0): (12, 1),
        (2048, 0): (11, 0), (8191, 0): (14, 1), (8192, 0): (13, 0), (16384, 0): (14, 0),
        (8193, 0): (0, 1), (3, -1): (2, 0), (2001, -1): (11, 0), (16384, -1): (14, 0),
        (16385, -1): (0, 0), (5, 2): (0, 0), (2, 2): (1, 0), (100, 4096): (12, 0),
    }
    for (n, f), want in cases.items():
        c = emu.hilbert_emu_class(n, f, ctypes.byref(ex))
        assert (c, ex.value) == want, (n, f, c, ex.value)


@pytest.mark.parametrize("n_fft", [0, -1])
def test_emulator_matches_scipy(emu, n_fft):
    rng = np.random.default_rng(11 + n_fft)
    lengths = list(range(1, 70)) + [127, 128, 129, 255, 256, 257, 1000, 1023, 1024, 2001,
                                     2048, 4095, 4096, 5000, 8191, 8192]
    for N in lengths:
        rows = rng.standard_normal((3, N))  # odd row count: the last pair has one row
        h = _emu_h(emu, rows, n_fft)
        ref = np.imag(H.analytic(rows, n_fft))
        err = np.max(np.abs(h - ref)) / np.max(np.abs(rows))
        assert err <= 1e-13, (N, n_fft, err)


@pytest.mark.parametrize("L", [4, 16, 64, 2048, 16384])
def test_emulator_fixed_length(emu, L):
    rng = np.random.default_rng(L)
    for N in sorted({1, 2, 3, L // 2 + 1, L}):
        rows = rng.standard_normal((2, N))
        h = _emu_h(emu, rows, L)
        ref = np.imag(H.analytic(rows, L))
        assert np.max(np.abs(h - ref)) <= 1e-13 * np.max(np.abs(rows)), (L, N)


# ---- the C ABI's argument checks (return before any device work) ----------------------

def _call(**kw):
    from cuda_emd import _lib
    lib = _lib.load()
    a = dict(imfs=None, row_base=None, max_imfs=4, n_imfs=None, offsets=None, N=100, B=0,
             min_n=0, max_n=0, n_fft=0, dt=1.0)
    a.update(kw)
    rc = lib.emd_hilbert(a["imfs"], a["row_base"], a["max_imfs"], a["n_imfs"], a["offsets"],
                         a["N"], a["B"], a["min_n"], a["max_n"], a["n_fft"], a["dt"], None, None,
                         None, None)
    return rc, lib.emd_last_error().decode()


def test_abi_rejects_bad_arguments():
    from cuda_emd import _lib
    bad = [dict(B=-1), dict(N=-1), dict(n_fft=3), dict(n_fft=-2), dict(dt=0.0),
           dict(dt=float("nan")), dict(dt=float("inf")), dict(N=100, n_fft=64),
           dict(N=20000, n_fft=32768), dict(N=20000, n_fft=-1), dict(N=10000, n_fft=0),
           dict(max_imfs=0), dict(offsets=1, min_n=5, max_n=4), dict(offsets=1, min_n=-1),
           dict(offsets=1, min_n=1, max_n=9000, n_fft=0)]
    for kw in bad:
        rc, msg = _call(**kw)
        assert rc == _lib.EMD_ERR_VALUE and msg, kw
    # a power-of-two row of 16384 samples is a direct transform; an empty batch is a no-op
    assert _call(N=16384, n_fft=0)[0] == _lib.EMD_OK
    assert _call(N=1000, n_fft=1024)[0] == _lib.EMD_OK


def test_hilbert_spectrum_host_checks():
    torch = pytest.importorskip("torch")
    from cuda_emd.hilbert import hilbert_spectrum
    with pytest.raises(ValueError):
        hilbert_spectrum(torch.zeros(2, 3, 8, dtype=torch.float32), [1, 1])
    with pytest.raises(ValueError):
        hilbert_spectrum(torch.zeros(2, 3, 8, dtype=torch.float64), [1, 1], outputs=("nope",))
