"""The oracle's time-axis mode (the reference's use of T) pinned without the oracle's own
code: the normalisation's closed form, a knot list derived by hand from PyEMD's mirror
rules on a non-uniform axis, scipy's CubicSpline as an independent not-a-knot spline, and
an affine axis that normalises to the sample index exactly."""
import numpy as np
import pytest
from scipy.interpolate import CubicSpline

from oracle import pyemd_restated as O


def test_normalisation_closed_form():
    # (T - T[0]) / min(diff T), emd.py:83-88
    T = np.array([2.0, 2.5, 3.5, 4.0, 5.5])
    assert np.array_equal(O.normalize_time(T), np.array([0.0, 1.0, 3.0, 4.0, 7.0]))
    with pytest.raises(ValueError):
        O.normalize_time(np.array([0.0, 1.0, 1.0, 2.0]))
    with pytest.raises(ValueError):
        O.normalize_time(np.array([0.0, 2.0, 1.0]))


def test_mirror_on_a_non_uniform_axis_by_hand():
    # maxima at 2, 6, 9 and minima at 4, 8; the first extremum is a maximum and
    # S[0] > S[4], so the left side mirrors about the first maximum (lsym = 2) with maxima
    # 9, 6 and minima 8, 4; the last extremum is a maximum and S[-1] > S[8], so the right
    # side mirrors about it (rsym = 9) with maxima 6, 2 and minima 8, 4.  Positions
    # 2 T[sym] - T[idx]: left 2*3 - T[..], right 2*13 - T[..]; neither side needs the
    # re-mirror about the signal edge.
    S = np.array([0.0, 1, 3, 1, -2, 1, 4, 0, -3, 2, 1, 0.5])
    T = np.array([0.0, 1, 3, 4, 6, 7, 8, 10, 11, 13, 14, 15])  # already normalised
    imax, imin, _ = O.find_extrema_simple(S)
    assert list(imax) == [2, 6, 9] and list(imin) == [4, 8]
    mk, nk = O.prepare_points_simple(S, imax, imin, 2, T)
    assert np.array_equal(mk[0], [-7, -2, 3, 8, 13, 18, 23])
    assert np.array_equal(mk[1], [2, 4, 3, 4, 2, 4, 3])
    assert np.array_equal(nk[0], [-5, 0, 6, 11, 15, 20])
    assert np.array_equal(nk[1], [-3, -2, -2, -3, -3, -2])


def test_envelope_is_the_not_a_knot_spline_at_the_axis():
    rng = np.random.default_rng(5)
    T = np.cumsum(rng.uniform(0.5, 1.5, 300))
    S = np.sin(2 * np.pi * np.arange(300) / 23.0) + 0.3 * rng.standard_normal(300)
    r = O.envelopes(S, T=T)
    Tn = O.normalize_time(T)
    for env, knots in ((r["max_env"], r["max_knots"]), (r["min_env"], r["min_knots"])):
        ref = CubicSpline(knots[0], knots[1], bc_type="not-a-knot")(Tn)
        assert np.max(np.abs(env - ref)) <= 1e-12 * np.max(np.abs(S))


def test_affine_axis_equals_the_index():
    from cuda_emd import synth
    x = synth.chirp_mix(3, 500, seed=2)
    for s in x:
        a = O.emd(s, max_imf=15)
        b = O.emd(s, max_imf=15, T=0.25 * np.arange(500) - 7.0)
        assert a.n_imfs == b.n_imfs and np.array_equal(a.imfs, b.imfs)
        assert np.array_equal(a.debug_n, b.debug_n)
