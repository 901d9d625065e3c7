"""Pins of the oracle (oracle/pyemd_restated.py) that do not come from the oracle's own
code: knot lists derived by hand for every branch of the nbsym=2 boundary mirror, stop
statistics that straddle each threshold in PyEMD's decision order, and the summation
order of the residue.

Mirror (reference kernels.py:382-503, PyEMD _prepare_points_simple).  Each case lists the
true extrema of a short signal and the knot positions/values that the branch table gives
for it, worked out by hand in the comments (T = arange(N), nbsym = 2):

* left branches, by the first extremum and the edge value:
  L1 first is a max, S[0] > S[first min]: lmax = ind_max[1:3] reversed, lmin = ind_min[0:2]
     reversed, mirrored about the first max;
  L2 first is a max, S[0] <= S[first min]: lmax = ind_max[0:2] reversed, lmin =
     [ind_min[0], 0], mirrored about sample 0;
  L3 first is a min, S[0] < S[first max]: lmax = ind_max[0:2] reversed, lmin = ind_min[1:3]
     reversed, mirrored about the first min;
  L4 first is a min, S[0] >= S[first max]: lmax = [ind_max[0], 0], lmin = ind_min[0:2]
     reversed, mirrored about sample 0;
* right branches R1-R4 the same way at the end (R2/R4 put sample N-1 first);
* a mirrored list that does not reach the edge is mirrored again about the edge;
* an empty list falls back to the whole (ascending) index array;
* equal adjacent knot positions keep one knot.
"""
import numpy as np
import pytest

from oracle import pyemd_restated as O

# (name, signal, max knots (pos, val), min knots (pos, val))
MIRROR_CASES = [
    # A: maxima 1,3,5 (1,2,3), minima 2,4,6 (-1,-2,-3); first is a max and S[0]=0.5 >
    # S[2]=-1 -> L1: lsym=1, lmax=[5,3] -> -3,-1; lmin=[4,2] -> -2,0.  Last is a min (6)
    # and S[7]=0 < S[5]=3 -> R1: rsym=6, rmax=[5,3] -> 7,9; rmin=[4,2] -> 8,10.
    ("A_L1_R1", [0.5, 1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 0.0],
     ([-3, -1, 1, 3, 5, 7, 9], [3, 2, 1, 2, 3, 3, 2]),
     ([-2, 0, 2, 4, 6, 8, 10], [-2, -1, -1, -2, -3, -2, -1])),
    # B: same extrema; S[0]=-1.5 <= S[2]=-1 -> L2: lsym=0, lmax=[3,1] -> -3,-1,
    # lmin=[2,0] -> -2,0 (values S2=-1, S0=-1.5).  S[7]=4 >= S[5]=3 -> R2: rsym=7,
    # rmax=[7,5] -> 7,9 (values 4,3); rmin=[6,4] -> 8,10.
    ("B_L2_R2", [-1.5, 1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 4.0],
     ([-3, -1, 1, 3, 5, 7, 9], [2, 1, 1, 2, 3, 4, 3]),
     ([-2, 0, 2, 4, 6, 8, 10], [-1, -1.5, -1, -2, -3, -3, -2])),
    # C: minima 1,3,5 (-1,-2,-3), maxima 2,4,6 (1,2,3); first is a min and S[0]=0.5 <
    # S[2]=1 -> L3: lsym=1, lmax=[4,2] -> -2,0; lmin=[5,3] -> -3,-1.  Last is a max (6)
    # and S[7]=0 > S[5]=-3 -> R3: rsym=6, rmax=[4,2] -> 8,10; rmin=[5,3] -> 7,9.
    ("C_L3_R3", [0.5, -1.0, 1.0, -2.0, 2.0, -3.0, 3.0, 0.0],
     ([-2, 0, 2, 4, 6, 8, 10], [2, 1, 1, 2, 3, 2, 1]),
     ([-3, -1, 1, 3, 5, 7, 9], [-3, -2, -1, -2, -3, -3, -2])),
    # D: same extrema; S[0]=1.5 >= S[2]=1 -> L4: lsym=0, lmax=[2,0] -> -2,0 (values 1,1.5);
    # lmin=[3,1] -> -3,-1.  S[7]=-4 <= S[5]=-3 -> R4: rsym=7, rmax=[6,4] -> 8,10;
    # rmin=[7,5] -> 7,9 (values -4,-3).
    ("D_L4_R4", [1.5, -1.0, 1.0, -2.0, 2.0, -3.0, 3.0, -4.0],
     ([-2, 0, 2, 4, 6, 8, 10], [1, 1.5, 1, 2, 3, 3, 2]),
     ([-3, -1, 1, 3, 5, 7, 9], [-2, -1, -1, -2, -3, -4, -3])),
    # E: maxima 4,6,8 (1,2,3), minima 5,7,9 (-1,-2,-3).  L1 about 4: lmin=[7,5] -> 1,3;
    # the first mirrored min (1) does not reach 0, so the left side is mirrored again
    # about 0 with lmax=[6,4] -> -6,-4 and lmin -> -7,-5.  R1 about 9: rmax=[8,6] -> 10,12,
    # rmin=[7,5] -> 11,13 (both reach 11).
    ("E_left_remirror", [0.0, 0.1, 0.2, 0.3, 1.0, -1.0, 2.0, -2.0, 3.0, -3.0, 0.5, 0.6],
     ([-6, -4, 4, 6, 8, 10, 12], [2, 1, 1, 2, 3, 3, 2]),
     ([-7, -5, 5, 7, 9, 11, 13], [-2, -1, -1, -2, -3, -2, -1])),
    # F: E reversed: minima 2,4,6 (-3,-2,-1), maxima 3,5,7 (3,2,1).  L3 about 2: lmax=[5,3]
    # -> -1,1; lmin=[6,4] -> -2,0.  R3 about 7: rmin=[6,4] -> 8,10 does not reach 11, so
    # the right side is mirrored again about 11: rmax=[7,5] -> 15,17, rmin -> 16,18.
    ("F_right_remirror", [0.6, 0.5, -3.0, 3.0, -2.0, 2.0, -1.0, 1.0, 0.3, 0.2, 0.1, 0.0],
     ([-1, 1, 3, 5, 7, 15, 17], [2, 3, 3, 2, 1, 1, 2]),
     ([-2, 0, 2, 4, 6, 16, 18], [-1, -2, -3, -2, -1, -1, -2])),
    # G: maxima 1 (2), 5 (3); one minimum 3 (-1).  L1 about 1: lmax=[5] -> -3; lmin=[3] ->
    # -1.  Last is a max and S[6]=0 > S[3]=-1 -> R3 about 5: rmax=[1] -> 9; rmin=[3] -> 7.
    # The minima envelope has three knots (PyEMD's natural 3-point spline).
    ("G_three_knots", [0.0, 2.0, 1.0, -1.0, 1.0, 3.0, 0.0],
     ([-3, 1, 5, 9], [3, 2, 3, 2]),
     ([-1, 3, 7], [-1, -1, -1])),
]


@pytest.mark.parametrize("case", MIRROR_CASES, ids=[c[0] for c in MIRROR_CASES])
def test_mirror_branches_by_hand(case):
    name, S, (mp, mv), (np_, nv) = case
    S = np.asarray(S, dtype=np.float64)
    imax, imin, _ = O.find_extrema_simple(S)
    mk, nk = O.prepare_points_simple(S, imax, imin, 2)
    assert np.array_equal(mk[0], mp) and np.array_equal(mk[1], mv), (name, mk)
    assert np.array_equal(nk[0], np_) and np.array_equal(nk[1], nv), (name, nk)


def test_mirror_fallback_and_both_remirrors_by_hand():
    """Index lists given directly (not alternating), so the empty-list fallback runs.
    S[0]=0.5 > S[4]=-1 with the first max (1) before the first min (4) -> L1 about 1:
    lmax = ind_max[1:1] is empty -> falls back to ind_max = [1], lmin = [6,4].  The
    mirrored max 2*1-1 = 1 does not reach 0 -> mirrored again about 0 with lmax = [1]:
    -1; lmin -> -6,-4.  Last is a min (6), S[9]=0.5 < S[1]=2 -> R1 about 6: rmax = [1],
    rmin = ind_min[0:1] = [4] -> 8, which does not reach 9 -> mirrored again about 9 with
    rmin = [6,4]: 12,14, and rmax: 17."""
    S = np.array([0.5, 2.0, 1.0, 0.0, -1.0, 0.0, -2.0, 0.0, 1.0, 0.5])
    mk, nk = O.prepare_points_simple(S, np.array([1]), np.array([4, 6]), 2)
    assert np.array_equal(mk[0], [-1, 1, 17]) and np.array_equal(mk[1], [2, 2, 2])
    assert np.array_equal(nk[0], [-6, -4, 4, 6, 12, 14])
    assert np.array_equal(nk[1], [-2, -1, -1, -2, -2, -1])


def test_mirror_duplicate_position_keeps_one_knot():
    """A minimum listed at sample 0 (possible only for a constructed index list): L4 puts
    sample 0 at the end of the mirrored minima (2*0 - 0 = 0), right before the true
    minimum at 0; the two equal positions become one knot."""
    S = np.array([3.0, 1.0, 2.0, 0.5, -1.0, 2.5, 0.0, 1.0])
    # first is a min (0 < 2), S[0]=3 >= S[2]=2 -> L4 about 0: lmax = [2, 0] -> -2, 0;
    # lmin = ind_min[0:2] reversed = [4, 0] -> -4, 0 (then the true 0: one knot at 0)
    mk, nk = O.prepare_points_simple(S, np.array([2, 5]), np.array([0, 4]), 2)
    assert list(nk[0][:3]) == [-4.0, 0.0, 4.0]
    assert list(nk[1][:3]) == [-1.0, 3.0, -1.0]
    assert list(mk[0][:4]) == [-2.0, 0.0, 2.0, 5.0]


# ---------------------------------------------------------------------------------------
# Stop statistics, PyEMD check_imf order (reference kernels.py:814-820): knot signs, then
# total power < 1e-10, then svar < 0.001, std < 0.2, energy ratio < 0.2.  Each pair of
# cases straddles one threshold while the earlier tests pass and the later ones fail, so
# the outcome is that test's alone; the last group shows the order itself matters.
# ---------------------------------------------------------------------------------------
GOOD = (np.array([[0.0, 1.0], [1.0, 1.0]]), np.array([[0.0, 1.0], [-1.0, -1.0]]))


def _pair(delta, scale=1.0, n=100, old_range=2.0):
    """imf_old spans old_range; imf_new = imf_old + delta at one sample."""
    old = scale * np.linspace(-old_range / 2, old_range / 2, n)
    old[old == 0] = 0.01
    new = old.copy()
    new[n // 3] += delta
    return new, old


def test_check_imf_sign_test_comes_first():
    new, old = _pair(1e-6)                       # svar tiny: would pass
    bad_max = (np.array([[0.0, 1.0], [-1.0, 1.0]]), GOOD[1])
    bad_min = (GOOD[0], np.array([[0.0, 1.0], [-1.0, 0.5]]))
    assert O.check_imf(new, old, *GOOD)
    assert not O.check_imf(new, old, *bad_max)
    assert not O.check_imf(new, old, *bad_min)


def test_check_imf_total_power_floor():
    # sum(new^2) just below / above 1e-10 with svar passing (diff 0)
    x = np.full(10, np.sqrt(1e-11) * 0.999)      # sum = 0.998e-10
    assert not O.check_imf(x, x, *GOOD)
    y = np.full(10, np.sqrt(1e-11) * 1.001)      # sum = 1.002e-10
    assert O.check_imf(y, y, *GOOD)


@pytest.mark.parametrize("below", [True, False])
def test_check_imf_svar_threshold(below):
    """svar = sum(d^2) / range(old) straddles 0.001 while std and the energy ratio fail:
    old spans 0.005 (sum(old^2) ~ 1.25e-5) and one sample moves by d with
    d^2 = 0.001 * 0.005 * (1 -+ 1e-6): energy ~ 0.4, (d/new)^2 ~ 0.92."""
    old = np.array([-0.0025, 0.0025, 1e-4, 1e-4, 1e-4, 1e-4])
    d = np.sqrt(0.001 * 0.005 * (1 - 1e-6 if below else 1 + 1e-6))
    new = old.copy()
    new[2] += d
    diff = new - old
    assert np.sum(diff * diff) / np.sum(old * old) >= 0.2
    assert np.sum((diff / new) ** 2) >= 0.2
    assert (np.sum(diff * diff) / (old.max() - old.min()) < 0.001) == below
    assert O.check_imf(new, old, *GOOD) == below


@pytest.mark.parametrize("below", [True, False])
def test_check_imf_std_threshold(below):
    """std = sum((d/new)^2) straddles 0.2 while svar and the energy ratio fail: three
    samples near c = 1e-4 (range 1e-6), one moved so that d/new = r with
    r^2 = 0.2 (1 -+ 1e-9): svar ~ 6.5, energy ~ 0.217."""
    c = 1e-4
    old = np.array([c, c * 1.01, c])
    r = np.sqrt(0.2 * (1 - 1e-9 if below else 1 + 1e-9))
    new = old.copy()
    new[0] = old[0] / (1 - r)
    diff = new - old
    assert np.sum(diff * diff) / (old.max() - old.min()) >= 0.001
    assert np.sum(diff * diff) / np.sum(old * old) >= 0.2
    assert (np.sum((diff / new) ** 2) < 0.2) == below
    assert O.check_imf(new, old, *GOOD) == below


@pytest.mark.parametrize("below", [True, False])
def test_check_imf_energy_ratio_threshold(below):
    # svar and std fail, energy ratio = sum(diff^2)/sum(old^2) straddles 0.2
    n = 8
    old = np.array([1e-4, 1e-4 + 1e-6, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4, 1e-4])
    target = 0.2 * (1 - 1e-9 if below else 1 + 1e-9) * np.sum(old * old)
    new = old.copy()
    new[2] = 1e-4 + np.sqrt(target)              # one sample moves; diff/new ~ 0.3 -> std
    diff = new - old
    assert np.sum(diff * diff) / (old.max() - old.min()) >= 0.001
    assert np.sum((diff / new) ** 2) >= 0.2
    er = np.sum(diff * diff) / np.sum(old * old)
    assert (er < 0.2) == below
    assert O.check_imf(new, old, *GOOD) == below


def test_check_imf_order_matters():
    """svar passes but total power is below 1e-10: PyEMD answers False (power first);
    the sign test overrides everything."""
    x = np.full(10, 1e-7)
    y = x.copy()
    y[3] += 1e-9                                  # svar minute
    assert not O.check_imf(y, x, *GOOD)
    big_new, big_old = _pair(1e-6)
    assert O.check_imf(big_new, big_old, *GOOD)


@pytest.mark.parametrize("kind", ["range", "power"])
def test_end_condition_thresholds(kind):
    """PyEMD end_condition (reference kernels.py:1037): stop when range(S - sum IMF) <
    0.001, or sum|S - sum IMF| < 0.005."""
    n = 50
    S = np.linspace(0.0, 1.0, n)
    if kind == "range":
        for r, want in ((0.001 * (1 - 1e-9), True), (0.001 * (1 + 1e-9), False)):
            imf = S - np.linspace(0.0, r, n) - 5.0   # residue spans exactly r
            # the total |residue| is large, so only the range test can stop
            assert O.end_condition(S, imf[None, :]) == want
    else:
        for t, want in ((0.005 * (1 - 1e-9), True), (0.005 * (1 + 1e-9), False)):
            res = np.zeros(n)
            res[0], res[1] = t / 2, -t / 2             # range t >= 0.001, sum|r| = t
            imf = S - res
            assert O.end_condition(S, imf[None, :]) == want


# ---------------------------------------------------------------------------------------
# Residue order.  PyEMD forms residue = S - np.sum(IMF[:k], axis=0); NumPy reduces a C
# array over axis 0 row by row (acc = r0; acc += r1; ...), not pairwise, and the kernel's
# running sum adds rows in the same order.  These rows tell the two orders apart:
# sequential ((1e16 + 1) - 1e16) + 1 = 1; pairwise (1e16 + 1) + (-1e16 + 1) = 0.
# ---------------------------------------------------------------------------------------
def test_residue_sum_is_sequential_over_rows():
    rows = np.array([[1e16, 3.0], [1.0, 0.1], [-1e16, -3.0], [1.0, 0.2]])
    seq = rows[0].copy()
    for r in rows[1:]:
        seq = seq + r
    pair = (rows[0] + rows[1]) + (rows[2] + rows[3])
    assert seq[0] == 1.0 and pair[0] == 0.0
    assert np.array_equal(np.sum(rows, axis=0), seq)


def test_oracle_residue_is_signal_minus_sequential_sum():
    """On a real decomposition the appended residue row is S minus the sequential sum of
    the IMF rows, bit for bit (and the pairwise order gives other bits on some sample, so
    the check discriminates)."""
    from cuda_emd import synth
    x = synth.seismic_traces(1, 700, seed=9)[0]
    r = O.emd(x, max_imf=15)
    assert r.n_imfs >= 4
    rows = r.imfs[:-1]
    seq = rows[0].copy()
    for q in rows[1:]:
        seq = seq + q
    assert np.array_equal(r.imfs[-1], x - seq)
    k = len(rows)
    half = (k + 1) // 2
    pair = rows[:half].sum(axis=0) + rows[half:].sum(axis=0)
    assert not np.array_equal(x - pair, x - seq)


# ---------------------------------------------------------------------------------------
# Reference mode (the reference kernel's own rules), each on a constructed case whose
# answer follows from the rule as written in reference kernels.py, and differs from
# PyEMD's where the two disagree.
# ---------------------------------------------------------------------------------------
def test_reference_zero_runs_counted_when_they_end():
    """Serial path (a zero inside the trace triggers it): each run of exact zeros counts
    once when it ends (kernels.py:126-143), so the trailing run is not counted; PyEMD
    counts every run.  S: +, 0, 0, -, +, 0 -> runs [1,2] (ends at 3), [5] (never ends);
    sign changes: (3,4)."""
    S = np.array([1.0, 0.0, 0.0, -1.0, 2.0, 0.0])
    assert O.find_extrema_reference(S)[2] == 2
    assert O.find_extrema_simple(S)[2] == 3


def test_reference_parallel_path_ignores_edge_zeros():
    """No zero or flat sample inside [1, N-2]: the parallel path counts only sign changes
    with a nonzero left sample (kernels.py:272-275), so zeros at sample 0 and N-1 do not
    count; PyEMD counts them."""
    S = np.array([0.0, 1.0, -1.0, 2.0, 0.0])
    assert O.find_extrema_reference(S)[2] == 2
    assert O.find_extrema_simple(S)[2] == 4


def test_reference_plateau_midpoint_floors():
    """A plateau over samples 2..5 between a rise and a fall: the reference puts the maximum
    at floor((2+5)/2) = 3 (kernels.py:191), PyEMD at round(3.5) = 4 (half to even)."""
    S = np.array([0.0, 1.0, 2.0, 2.0, 2.0, 2.0, 1.0, 0.0])
    assert list(O.find_extrema_reference(S)[0]) == [3]
    assert list(O.find_extrema_simple(S)[0]) == [4]
    # a plateau over 2..4: floor(3) = round(3) = 3 in both
    S = np.array([0.0, 1.0, 2.0, 2.0, 2.0, 1.0, 0.0])
    assert list(O.find_extrema_reference(S)[0]) == list(O.find_extrema_simple(S)[0]) == [3]


def test_reference_plateau_at_start_is_not_seen():
    """The serial scan finds a plateau at its first sample i >= 1 with a sloped sample
    before it, so a plateau starting at sample 0 is never an extremum; a plateau reaching
    the end is dropped in both."""
    S = np.array([1.0, 1.0, 0.0, 2.0, 1.0, 1.0])
    mx, mn, _ = O.find_extrema_reference(S)
    assert list(mx) == [3] and list(mn) == [2]


def test_reference_three_knot_parabola():
    """3 knots: the reference evaluates the interpolating parabola (kernels.py:711-718):
    exact at the knots, constant second difference; PyEMD's natural cubic differs."""
    ext = np.array([[0.0, 3.0, 8.0], [1.0, -2.0, 4.0]])
    T = np.arange(9, dtype=np.float64)
    q = O.spline_points(T, ext, mode="reference")
    assert np.allclose(q[[0, 3, 8]], [1.0, -2.0, 4.0], rtol=0, atol=1e-14)
    assert np.allclose(np.diff(q, 2), np.diff(q, 2)[0], rtol=0, atol=1e-12)
    c = O.spline_points(T, ext)
    assert np.max(np.abs(c - q)) > 1e-3


def test_reference_running_residue():
    """Reference mode: the appended residue is S minus the IMF rows subtracted one at a
    time (kernels.py:996), bit for bit; PyEMD's is S minus their sum."""
    from cuda_emd import synth
    x = synth.seismic_traces(1, 700, seed=9)[0]
    r = O.emd(x, max_imf=15, mode="reference")
    res = x.copy()
    for q in r.imfs[:-1]:
        res = res - q
    assert np.array_equal(r.imfs[-1], res)
    p = O.emd(x, max_imf=15)
    assert p.n_imfs == r.n_imfs and np.array_equal(p.debug_n, r.debug_n)
    assert not np.array_equal(p.imfs[-1], r.imfs[-1])
    assert np.max(np.abs(p.imfs - r.imfs)) < 1e-12 * np.max(np.abs(x))
