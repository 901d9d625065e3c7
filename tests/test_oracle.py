"""Pins the CPU oracle (oracle/pyemd_restated.py) against facts it does not restate:
closed forms, an independent scipy routine, brute force over every short pattern, and
invariants.  Runs on the CPU."""
import itertools

import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import pyemd_restated as O


# ---------------------------------------------------------------------------------------
# splines
# ---------------------------------------------------------------------------------------
def test_notaknot_reproduces_cubics():
    """Not-a-knot reproduces any cubic polynomial exactly (a natural spline does not:
    its end second derivatives are forced to 0)."""
    rng = np.random.default_rng(1)
    for n in (4, 5, 9, 40):
        x = np.sort(rng.choice(np.arange(-20, 120), n, replace=False)).astype(float)
        c = rng.standard_normal(4)
        p = lambda t: c[0] + c[1] * t + c[2] * t ** 2 + c[3] * t ** 3
        T = np.arange(0, 100, dtype=float)
        got = O.spline_points(T, np.array([x, p(x)]))
        t = T[(T >= x[0]) & (T <= x[-1])]
        assert np.allclose(got, p(t), rtol=1e-9, atol=1e-9 * np.max(np.abs(p(t))))


def test_notaknot_matches_scipy_cubicspline():
    rng = np.random.default_rng(2)
    for _ in range(50):
        n = int(rng.integers(4, 60))
        x = np.cumsum(rng.integers(1, 9, n)).astype(float) - 5
        y = rng.standard_normal(n)
        T = np.arange(0, x[-1] + 1, dtype=float)
        got = O.spline_points(T, np.array([x, y]))
        t = T[(T >= x[0]) & (T <= x[-1])]
        ref = CubicSpline(x, y, bc_type="not-a-knot")(t)
        assert np.max(np.abs(got - ref)) <= 1e-11 * max(1.0, np.max(np.abs(y)))


def test_three_point_spline_is_natural_cubic():
    """Interpolates the knots, S'' = 0 at both ends, C1 and C2 at the middle knot."""
    rng = np.random.default_rng(3)
    for _ in range(30):
        x = np.sort(rng.choice(np.arange(-30, 60), 3, replace=False)).astype(float)
        y = rng.standard_normal(3)
        t = np.linspace(x[0], x[2], 2001)
        q = O.cubic_spline_3pts(x, y, t)
        assert np.allclose(O.cubic_spline_3pts(x, y, x), y, atol=1e-12)
        ref = CubicSpline(x, y, bc_type="natural")(t)
        assert np.max(np.abs(q - ref)) <= 1e-12 * max(1.0, np.max(np.abs(y)))


def test_three_point_collinear_is_line():
    x = np.array([-3.0, 4.0, 10.0])
    y = 0.5 * x + 2.0
    t = np.arange(-3, 11, dtype=float)
    assert np.allclose(O.cubic_spline_3pts(x, y, t), 0.5 * t + 2.0, atol=1e-13)


# ---------------------------------------------------------------------------------------
# extrema / zero crossings: brute force from sample runs (S-space), independent of the
# vectorised d-space implementation
# ---------------------------------------------------------------------------------------
def _rhe(s):
    return int(np.round(s / 2.0))


def brute_extrema(S):
    N = len(S)
    imax, imin = [], []
    for i in range(1, N - 1):
        if S[i] > S[i - 1] and S[i] > S[i + 1]:
            imax.append(i)
        elif S[i] < S[i - 1] and S[i] < S[i + 1]:
            imin.append(i)
    # maximal runs of >= 2 equal samples [a, b]
    runs = []
    i = 0
    while i < N - 1:
        if S[i + 1] == S[i]:
            j = i + 1
            while j + 1 < N and S[j + 1] == S[i]:
                j += 1
            runs.append((i, j))
            i = j
        else:
            i += 1
    keep = []
    for k, (a, b) in enumerate(runs):
        if k == 0 and a == 1:
            continue               # PyEMD: first run starting at d-index 1 is dropped
        if b == N - 1:
            continue               # run reaching the end is dropped
        before = (S[a] - S[a - 1]) if a >= 1 else (S[N - 1] - S[N - 2])  # d[-1] wraparound
        after = S[b + 1] - S[b]
        if before > 0 and after < 0:
            keep.append(("max", _rhe(a + b)))
        elif not before > 0 and after > 0:
            keep.append(("min", _rhe(a + b)))
    imax += [m for t, m in keep if t == "max"]
    imin += [m for t, m in keep if t == "min"]
    nz = sum(1 for i in range(N - 1) if S[i] * S[i + 1] < 0)
    nz += sum(1 for i in range(N) if S[i] == 0 and (i == 0 or S[i - 1] != 0))
    return sorted(imax), sorted(imin), nz


def test_extrema_exhaustive_small_patterns():
    vals = (-1.0, 0.0, 1.0, 2.0)
    count = 0
    for n in range(1, 8):
        for pat in itertools.product(vals, repeat=n):
            S = np.array(pat)
            imax, imin, nz = O.find_extrema_simple(S)
            bmax, bmin, bnz = brute_extrema(S)
            assert list(imax) == bmax and list(imin) == bmin and nz == bnz, (pat, imax, imin, nz)
            count += 1
    assert count > 20000


def test_extrema_random_signals():
    rng = np.random.default_rng(4)
    for _ in range(200):
        n = int(rng.integers(3, 300))
        S = np.round(rng.standard_normal(n) * rng.choice([1, 3, 100]))  # many ties
        imax, imin, nz = O.find_extrema_simple(S)
        assert (list(imax), list(imin), nz) == brute_extrema(S)


# ---------------------------------------------------------------------------------------
# mirroring invariants
# ---------------------------------------------------------------------------------------
def test_mirror_invariants():
    rng = np.random.default_rng(5)
    for _ in range(300):
        N = int(rng.integers(8, 400))
        S = np.cumsum(rng.standard_normal(N)) if rng.uniform() < 0.5 else rng.standard_normal(N)
        imax, imin, _ = O.find_extrema_simple(S)
        if len(imax) + len(imin) <= 2 or not len(imax) or not len(imin):
            continue
        mk, nk = O.prepare_points_simple(S, imax, imin)
        for k, ind in ((mk, imax), (nk, imin)):
            assert k[0, 0] <= 0 and k[0, -1] >= N - 1          # knots cover the signal
            assert np.all(np.diff(k[0]) > 0)                    # strictly increasing
            assert np.all(k[0] == np.round(k[0]))               # integer positions
            inside = (k[0] >= 0) & (k[0] <= N - 1)
            # every true extremum is a knot with its own value
            assert set(ind.tolist()) <= set(k[0, inside].astype(int).tolist())
            pos = k[0].astype(int)
            for p, v in zip(pos, k[1]):
                if 0 <= p < N and p in set(ind.tolist()):
                    assert v == S[p]


def test_mirror_branch_table_example():
    """Hand-worked case: first extremum a maximum, S[0] above the first minimum ->
    lsym = first maximum, mirrored maxima are the 2nd/3rd maxima reflected about it."""
    S = np.array([0.5, 1.0, 0.0, 2.0, -1.0, 3.0, -2.0, 1.5, 0.2])
    imax, imin, _ = O.find_extrema_simple(S)
    assert list(imax) == [1, 3, 5, 7] and list(imin) == [2, 4, 6]
    mk, nk = O.prepare_points_simple(S, imax, imin)
    # left: lmax = [5, 3] reflected about 1 -> [-3, -1]; lmin = [4, 2] -> [-2, 0]
    assert list(mk[0, :2]) == [-3.0, -1.0] and list(mk[1, :2]) == [3.0, 2.0]
    assert list(nk[0, :2]) == [-2.0, 0.0] and list(nk[1, :2]) == [-1.0, 0.0]


# ---------------------------------------------------------------------------------------
# whole decomposition
# ---------------------------------------------------------------------------------------
def test_zero_signal_has_no_rows():
    r = O.emd(np.zeros(64), max_imf=15)
    assert r.n_imfs == 0 and list(r.debug_n) == [-1] * 16


def test_linear_trend_is_residue_only():
    x = 0.3 * np.arange(100) - 4.0
    r = O.emd(x, max_imf=15)
    assert r.n_imfs == 1 and np.array_equal(r.imfs[0], x)


def test_reconstruction_identity():
    rng = np.random.default_rng(6)
    for _ in range(20):
        x = np.cumsum(rng.standard_normal(int(rng.integers(50, 600))))
        r = O.emd(x, max_imf=15)
        assert np.max(np.abs(r.imfs.sum(axis=0) - x)) <= 1e-12 * np.max(np.abs(x))


def test_pure_tone_single_imf():
    n = np.arange(512)
    x = np.sin(2 * np.pi * n / 25.0 + 0.3)
    r = O.emd(x, max_imf=15)
    # one oscillatory IMF carrying the tone; any residue is tiny
    assert r.n_imfs >= 1
    assert np.max(np.abs(r.imfs[0] - x)) < 0.05


def test_max_iteration_and_cap_semantics():
    rng = np.random.default_rng(7)
    x = rng.standard_normal(300)
    r = O.emd(x, max_imf=15, max_iteration=2)
    assert r.overflow == 1 and r.debug_n[0] == 2 and r.debug_extno[0] == -1
    r = O.emd(x, max_imf=2, debug_slots=3)
    assert r.n_imfs <= 3


def test_max_extrema_clamp_sets_overflow():
    rng = np.random.default_rng(8)
    x = rng.standard_normal(400)
    assert O.emd(x, max_imf=15, max_extrema=10).overflow == 1
    assert O.emd(x, max_imf=15).overflow == 0


def test_near_tie_logging():
    n = np.arange(257)
    r = O.emd(np.sin(2 * np.pi * n / 32.0), max_imf=15)
    assert any(t.what == "sample_sign" for t in r.near_ties)


def test_envelopes_cover_signal():
    rng = np.random.default_rng(9)
    x = rng.standard_normal(300)
    e = O.envelopes(x)
    assert e["degenerate"] == 0
    imax, imin, _ = O.find_extrema_simple(x)
    assert np.allclose(e["max_env"][imax], x[imax]) and np.allclose(e["min_env"][imin], x[imin])
