"""CPU oracle for the batched EMD sift: a NumPy/SciPy restatement of PyEMD ``EMD()``
at its defaults (extrema_detection='simple', spline_kind='cubic', nbsym=2, FIXE=FIXE_H=0,
MAX_ITERATION=1000, thresholds energy_ratio 0.2 / std 0.2 / svar 0.001 / total_power 0.005 /
range 0.001).

TEST INFRASTRUCTURE ONLY. Nothing in the product imports this module: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and only as the
checker / the reported CPU baseline, never as the measured GPU path.

Provenance.  The reference (SeismicFlow/emd-cuda) states that its kernel reproduces
PyEMD ``EMD()`` defaults (kernels.py:48-53, README.md:5-6, 74-93).  PyEMD itself
(package ``EMD-signal``, Apache-2.0) is not installed in this image and its version is
unpinned by the reference, so the functions below restate its published algorithm; each
function names the PyEMD routine it restates and the reference lines that implement the
same step.  Where the reference and PyEMD differ (3-knot spline, plateau midpoint rounding,
trailing zero runs, residue summation order, IMF cap) this module follows PyEMD; the two
reference-only rules that have no PyEMD counterpart are available as options:

* ``max_extrema`` -- the reference's shared-memory capacity rule: counts are clamped to
  ``max_extrema`` per side and ``overflow`` is set when a count reaches it
  (kernels.py:196-203, 318-344).  ``None`` = PyEMD (no cap).
* ``max_imf`` -- PyEMD's own ``max_imf`` argument; the reference's loop bound
  ``imf_count < MAX_IMFS - 1`` (kernels.py:894) is ``max_imf = MAX_IMFS - 1``.
* a signal with ``extNo > 2`` but no maximum or no minimum is treated as a trend, as the
  reference does (kernels.py:934); PyEMD raises IndexError there.

``mode='reference'`` switches four rules to the reference kernel's own (for measuring how
often the two semantics part ways; the HIP kernels implement PyEMD's):

* extrema and zero crossings by the reference's scan (find_extrema_reference): its
  parallel path when no sample in [1, N-2] is zero or flat, else its serial path
  (kernels.py:115-204, 231-349);
* a 3-knot envelope is the interpolating parabola (kernels.py:711-718);
* the residue is updated by subtraction, residue -= imf, and the end condition, the trend
  drop and the appended residue use that running residue (kernels.py:994-1075).

Near ties: every stop statistic within 1e-12 (relative) of its threshold, and every
sign test (a sample or first difference within 1e-12 * max|imf| of zero, after the raw
input) is logged in ``EMDResult.near_ties``; only there may rounding-level differences
between two correct implementations change a count.

Pinning (see tests/test_oracle.py): the spline against scipy ``CubicSpline``
(not-a-knot) and closed forms; the 3-point spline against the natural-spline
conditions; extrema/zero crossings against a brute-force definition on every sign /
plateau pattern of short signals; the mirror against its branch table and the
"knots cover [T0, T_{N-1}]" invariant; whole-EMD special cases (zero signal, linear
trend, pure tone) and the reconstruction identity.  The plateau edge-run rules
(``debs[0] == 1`` drop, ``d[-1]`` wraparound) and PyEMD's exact residue summation order
are restated from PyEMD's source as recalled; nothing in this image can execute PyEMD
to confirm them, so those two rules are "parity unpinned" beyond the brute-force
definition written in the test from the same recollection.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy.interpolate import interp1d

ENERGY_RATIO_THR = 0.2
STD_THR = 0.2
SVAR_THR = 0.001
TOTAL_POWER_THR = 0.005
RANGE_THR = 0.001
NEAR_TIE_REL = 1e-12


# ---------------------------------------------------------------------------------------
# Extrema ("simple" mode)
# ---------------------------------------------------------------------------------------
def find_extrema_simple(S: np.ndarray):
    """PyEMD ``EMD._find_extrema_simple``: strict extrema by ``d1*d2 < 0``, plateau
    extrema at the rounded run midpoint, zero crossings with zero runs collapsed.

    Reference counterpart: find_extrema_serial / find_extrema_parallel
    (kernels.py:115-204, 231-349).  Returns (indmax, indmin, nzm) with int64 indices.
    """
    S = np.asarray(S, dtype=np.float64)
    n = S.size
    if n < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), int(np.count_nonzero(S == 0))
    S1, S2 = S[:-1], S[1:]
    indzer = np.nonzero(S1 * S2 < 0)[0]
    if np.any(S == 0):
        indz = np.nonzero(S == 0)[0]
        if np.any(np.diff(indz) == 1):
            zer = S == 0
            dz = np.diff(np.append(np.append(0, zer), 0))
            debz = np.nonzero(dz == 1)[0]
            finz = np.nonzero(dz == -1)[0] - 1
            indz = np.round((debz + finz) / 2.0)
        indzer = np.sort(np.append(indzer, indz))

    d = np.diff(S)
    d1, d2 = d[:-1], d[1:]
    indmin = np.nonzero(np.r_[d1 * d2 < 0] & np.r_[d1 < 0])[0] + 1
    indmax = np.nonzero(np.r_[d1 * d2 < 0] & np.r_[d1 > 0])[0] + 1

    if np.any(d == 0):
        imax, imin = [], []
        bad = d == 0
        dd = np.diff(np.append(np.append(0, bad), 0))
        debs = np.nonzero(dd == 1)[0]
        fins = np.nonzero(dd == -1)[0]
        if debs[0] == 1:
            if len(debs) > 1:
                debs, fins = debs[1:], fins[1:]
            else:
                debs, fins = [], []
        if len(debs) > 0:
            if fins[-1] == len(S) - 1:
                if len(debs) > 1:
                    debs, fins = debs[:-1], fins[:-1]
                else:
                    debs, fins = [], []
        for k in range(len(debs)):
            if d[debs[k] - 1] > 0:
                if d[fins[k]] < 0:
                    imax.append(np.round((fins[k] + debs[k]) / 2.0))
            else:
                if d[fins[k]] > 0:
                    imin.append(np.round((fins[k] + debs[k]) / 2.0))
        if imax:
            indmax = np.sort(np.append(indmax, np.asarray(imax, np.int64)))
        if imin:
            indmin = np.sort(np.append(indmin, np.asarray(imin, np.int64)))

    return indmax.astype(np.int64), indmin.astype(np.int64), int(len(indzer))


def find_extrema_reference(S: np.ndarray):
    """The reference kernel's extrema and zero-crossing rules (reference mode).

    Parallel path (kernels.py:231-349), used when no sample i in [1, N-2] has S[i] == 0,
    S[i] == S[i-1] or S[i+1] == S[i]: strict extrema by d1*d2 < 0, zero crossings =
    #(S[i] != 0 and S[i]*S[i+1] < 0), so a zero at sample 0 or N-1 is not counted.
    Otherwise the serial path (kernels.py:115-204): a run of exact zeros counts once when
    it ends (a run reaching the end is never counted); a plateau is found at its first
    sample i >= 1 (so one starting at sample 0 is never seen), dropped when it reaches the
    end, and its extremum sits at floor((first + last) / 2).
    Returns (indmax, indmin, nzm)."""
    S = np.asarray(S, dtype=np.float64)
    N = S.size
    if N < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), 0
    inner = np.arange(1, N - 1)
    d = np.diff(S)
    trigger = np.any((S[inner] == 0) | (d[inner - 1] == 0) | (d[inner] == 0)) if N > 2 else False
    if not trigger:
        d1, d2 = d[:-1], d[1:]
        ext = d1 * d2 < 0
        indmax = np.nonzero(ext & (d1 > 0))[0] + 1
        indmin = np.nonzero(ext & (d1 < 0))[0] + 1
        nz = int(np.count_nonzero((S[:-1] != 0) & (S[:-1] * S[1:] < 0)))
        return indmax.astype(np.int64), indmin.astype(np.int64), nz
    nz = 0
    in_run = False
    for i in range(N):
        z = S[i] == 0.0
        if z and not in_run:
            in_run = True
        if not z and in_run:
            nz += 1
            in_run = False
        if not z and i + 1 < N and S[i] * S[i + 1] < 0.0:
            nz += 1
    imax, imin = [], []
    i = 1
    while i < N - 1:
        dprev = S[i] - S[i - 1]
        dnext = S[i + 1] - S[i]
        if dprev * dnext < 0.0:
            (imin if dprev < 0.0 else imax).append(i)
        elif dnext == 0.0 and dprev != 0.0:
            j = i
            while j + 1 < N and S[j + 1] == S[i]:
                j += 1
            if j + 1 < N:
                d_after = S[j + 1] - S[j]
                mid = (i + j) // 2
                if dprev > 0.0 and d_after < 0.0:
                    imax.append(mid)
                elif dprev < 0.0 and d_after > 0.0:
                    imin.append(mid)
            i = j
        i += 1
    return np.asarray(imax, np.int64), np.asarray(imin, np.int64), nz


# ---------------------------------------------------------------------------------------
# Boundary mirroring
# ---------------------------------------------------------------------------------------
def normalize_time(T) -> np.ndarray:
    """The reference's time normalisation (emd.py:83-88, PyEMD ``_normalize_time``):
    ``(T - T[0]) / min(diff T)``.  T must increase strictly (the reference divides by the
    smallest step, and its mirror and interval search assume an increasing axis)."""
    T = np.asarray(T, dtype=np.float64)
    if T.ndim != 1:
        raise ValueError("T must be one-dimensional")
    if T.size > 1:
        d = np.diff(T)
        if not np.all(d > 0):
            raise ValueError("T must increase strictly")
        return (T - T[0]) / np.min(d)
    return T - T[0] if T.size else T


def prepare_points_simple(S: np.ndarray, ind_max: np.ndarray, ind_min: np.ndarray, nbsym: int = 2,
                          T: np.ndarray | None = None):
    """PyEMD ``EMD._prepare_points_simple``: mirror ``nbsym`` extrema about the chosen
    symmetry points, re-mirror about the signal edge when the mirrored points do not reach
    it, then drop adjacent knots with equal positions.  T: the (normalised) time axis;
    None = arange(N), PyEMD's simple mode.

    Reference counterpart: mirror_one_side (kernels.py:360-558), which places knots at
    T[idx] and mirrored knots at 2 T[sym] - T[idx] (kernels.py:471-472, 528-552).  Returns
    (max_knots [2,k], min_knots [2,k]) rows (position, value).  The PyEMD "Left/Right
    edge BUG" exceptions are unreachable for an increasing T with T[0] = 0; they are kept.
    """
    N = S.size
    T = np.arange(N, dtype=np.float64) if T is None else np.asarray(T, dtype=np.float64)
    ind_min = np.asarray(ind_min, dtype=int)
    ind_max = np.asarray(ind_max, dtype=int)
    end_min, end_max = len(ind_min), len(ind_max)

    if ind_max[0] < ind_min[0]:
        if S[0] > S[ind_min[0]]:
            lmax = ind_max[1:min(end_max, nbsym + 1)][::-1]
            lmin = ind_min[0:min(end_min, nbsym + 0)][::-1]
            lsym = ind_max[0]
        else:
            lmax = ind_max[0:min(end_max, nbsym)][::-1]
            lmin = np.append(ind_min[0:min(end_min, nbsym - 1)][::-1], 0)
            lsym = 0
    else:
        if S[0] < S[ind_max[0]]:
            lmax = ind_max[0:min(end_max, nbsym + 0)][::-1]
            lmin = ind_min[1:min(end_min, nbsym + 1)][::-1]
            lsym = ind_min[0]
        else:
            lmax = np.append(ind_max[0:min(end_max, nbsym - 1)][::-1], 0)
            lmin = ind_min[0:min(end_min, nbsym)][::-1]
            lsym = 0

    if ind_max[-1] < ind_min[-1]:
        if S[-1] < S[ind_max[-1]]:
            rmax = ind_max[max(end_max - nbsym, 0):][::-1]
            rmin = ind_min[max(end_min - nbsym - 1, 0):-1][::-1]
            rsym = ind_min[-1]
        else:
            rmax = np.append(ind_max[max(end_max - nbsym + 1, 0):], N - 1)[::-1]
            rmin = ind_min[max(end_min - nbsym, 0):][::-1]
            rsym = N - 1
    else:
        if S[-1] > S[ind_min[-1]]:
            rmax = ind_max[max(end_max - nbsym - 1, 0):-1][::-1]
            rmin = ind_min[max(end_min - nbsym, 0):][::-1]
            rsym = ind_max[-1]
        else:
            rmax = ind_max[max(end_max - nbsym, 0):][::-1]
            rmin = np.append(ind_min[max(end_min - nbsym + 1, 0):], N - 1)[::-1]
            rsym = N - 1

    if not lmin.size:
        lmin = ind_min
    if not rmin.size:
        rmin = ind_min
    if not lmax.size:
        lmax = ind_max
    if not rmax.size:
        rmax = ind_max

    tlmin = 2 * T[lsym] - T[lmin]
    tlmax = 2 * T[lsym] - T[lmax]
    trmin = 2 * T[rsym] - T[rmin]
    trmax = 2 * T[rsym] - T[rmax]

    if tlmin[0] > T[0] or tlmax[0] > T[0]:
        if lsym == ind_max[0]:
            lmax = ind_max[0:min(end_max, nbsym)][::-1]
        else:
            lmin = ind_min[0:min(end_min, nbsym)][::-1]
        if lsym == 0:
            raise RuntimeError("Left edge BUG")
        lsym = 0
        tlmin = 2 * T[lsym] - T[lmin]
        tlmax = 2 * T[lsym] - T[lmax]

    if trmin[-1] < T[-1] or trmax[-1] < T[-1]:
        if rsym == ind_max[-1]:
            rmax = ind_max[max(end_max - nbsym, 0):][::-1]
        else:
            rmin = ind_min[max(end_min - nbsym, 0):][::-1]
        if rsym == N - 1:
            raise RuntimeError("Right edge BUG")
        rsym = N - 1
        trmin = 2 * T[rsym] - T[rmin]
        trmax = 2 * T[rsym] - T[rmax]

    tmin = np.append(tlmin, np.append(T[ind_min], trmin))
    tmax = np.append(tlmax, np.append(T[ind_max], trmax))
    zmin = np.append(S[lmin], np.append(S[ind_min], S[rmin]))
    zmax = np.append(S[lmax], np.append(S[ind_max], S[rmax]))

    max_extrema = np.array([tmax, zmax])
    min_extrema = np.array([tmin, zmin])
    max_dup_idx = np.where(max_extrema[0, 1:] == max_extrema[0, :-1])
    max_extrema = np.delete(max_extrema, max_dup_idx, axis=1)
    min_dup_idx = np.where(min_extrema[0, 1:] == min_extrema[0, :-1])
    min_extrema = np.delete(min_extrema, min_dup_idx, axis=1)
    return max_extrema, min_extrema


# ---------------------------------------------------------------------------------------
# Splines
# ---------------------------------------------------------------------------------------
def cubic_spline_3pts(x, y, t):
    """PyEMD ``splines.cubic_spline_3pts``: natural cubic spline through 3 knots in slope
    form (3x3 system solved by inversion), evaluated piecewise (t < x1 / t >= x1).

    The reference instead evaluates the interpolating parabola (kernels.py:568-572,
    711-718); this oracle follows PyEMD.
    """
    x0, x1, x2 = x
    y0, y1, y2 = y
    x1x0, x2x1 = x1 - x0, x2 - x1
    y1y0, y2y1 = y1 - y0, y2 - y1
    _x1x0, _x2x1 = 1.0 / x1x0, 1.0 / x2x1
    m11, m12, m13 = 2 * _x1x0, _x1x0, 0
    m21, m22, m23 = _x1x0, 2.0 * (_x1x0 + _x2x1), _x2x1
    m31, m32, m33 = 0, _x2x1, 2.0 * _x2x1
    v1 = 3 * y1y0 * _x1x0 * _x1x0
    v3 = 3 * y2y1 * _x2x1 * _x2x1
    v2 = v1 + v3
    M = np.array([[m11, m12, m13], [m21, m22, m23], [m31, m32, m33]])
    v = np.array([v1, v2, v3]).T
    k = np.array(np.linalg.inv(M).dot(v))
    a1 = k[0] * x1x0 - y1y0
    b1 = -k[1] * x1x0 + y1y0
    a2 = k[1] * x2x1 - y2y1
    b2 = -k[2] * x2x1 + y2y1
    t1 = (t[np.r_[t < x1]] - x0) / x1x0
    t2 = (t[np.r_[t >= x1]] - x1) / x2x1
    t11, t22 = 1.0 - t1, 1.0 - t2
    q1 = t11 * y0 + t1 * y1 + t1 * t11 * (a1 * t11 + b1 * t1)
    q2 = t22 * y1 + t2 * y2 + t2 * t22 * (a2 * t22 + b2 * t2)
    return np.append(q1, q2)


def spline_points(T: np.ndarray, extrema: np.ndarray, mode: str = "pyemd") -> np.ndarray:
    """PyEMD ``EMD.spline_points`` for spline_kind='cubic': evaluate on the samples of T
    inside [first knot, last knot]; more than 3 knots -> ``interp1d(kind='cubic')``
    (scipy ``make_interp_spline(k=3)``, not-a-knot), 3 knots -> ``cubic_spline_3pts``.
    Two knots (PyEMD cannot reach a valid result there) -> straight line, as the
    reference does (kernels.py:706-709).

    Reference counterpart: notaknot_cubic_spline_block (kernels.py:585-736).
    """
    t = T[np.r_[T >= extrema[0, 0]] & np.r_[T <= extrema[0, -1]]]
    nk = extrema.shape[1]
    if nk > 3:
        return interp1d(extrema[0], extrema[1], kind="cubic")(t)
    if nk == 3:
        if mode == "reference":
            # the reference's 3-knot envelope: the interpolating parabola, Lagrange form
            (x0, x1, x2), (y0, y1, y2) = extrema[0], extrema[1]
            l0 = (t - x1) * (t - x2) / ((x0 - x1) * (x0 - x2))
            l1 = (t - x0) * (t - x2) / ((x1 - x0) * (x1 - x2))
            l2 = (t - x0) * (t - x1) / ((x2 - x0) * (x2 - x1))
            return y0 * l0 + y1 * l1 + y2 * l2
        return cubic_spline_3pts(extrema[0], extrema[1], t)
    if nk == 2:
        x0, x1 = extrema[0]
        y0, y1 = extrema[1]
        return y0 + (y1 - y0) / (x1 - x0) * (t - x0)
    return np.full(t.shape, extrema[1, 0])


# ---------------------------------------------------------------------------------------
# Stop tests
# ---------------------------------------------------------------------------------------
@dataclass
class NearTie:
    imf: int
    sift: int
    what: str
    value: float
    threshold: float


def _near(value: float, thr: float) -> bool:
    return bool(np.isfinite(value)) and abs(value - thr) <= NEAR_TIE_REL * max(1.0, abs(thr))


def check_imf(imf_new, imf_old, eMax, eMin, ties=None, where=(0, 0)) -> bool:
    """PyEMD ``EMD.check_imf``: knot-sign test, then total power 1e-10, then scaled
    variance < svar_thr, then std < std_thr, then energy ratio < energy_ratio_thr.

    Reference counterpart: check_imf_block (kernels.py:742-821).
    """
    if np.any(eMax[1] < 0) or np.any(eMin[1] > 0):
        return False
    sq = np.sum(imf_new ** 2)
    if ties is not None and _near(sq, 1e-10):
        ties.append(NearTie(where[0], where[1], "total_power_1e-10", float(sq), 1e-10))
    if sq < 1e-10:
        return False
    imf_diff = imf_new - imf_old
    imf_diff_sqrd_sum = np.sum(imf_diff * imf_diff)
    svar = imf_diff_sqrd_sum / (max(imf_old) - min(imf_old))
    if ties is not None and _near(svar, SVAR_THR):
        ties.append(NearTie(where[0], where[1], "svar", float(svar), SVAR_THR))
    if svar < SVAR_THR:
        return True
    with np.errstate(divide="ignore", invalid="ignore"):
        std = np.sum((imf_diff / imf_new) ** 2)
    if ties is not None and _near(std, STD_THR):
        ties.append(NearTie(where[0], where[1], "std", float(std), STD_THR))
    if std < STD_THR:
        return True
    energy_ratio = imf_diff_sqrd_sum / np.sum(imf_old * imf_old)
    if ties is not None and _near(energy_ratio, ENERGY_RATIO_THR):
        ties.append(NearTie(where[0], where[1], "energy_ratio", float(energy_ratio), ENERGY_RATIO_THR))
    if energy_ratio < ENERGY_RATIO_THR:
        return True
    return False


def end_condition(S, IMF, ties=None, where=(0, 0)) -> bool:
    """PyEMD ``EMD.end_condition``: stop when range(S - sum IMF) < range_thr or
    sum|S - sum IMF| < total_power_thr.  Reference: kernels.py:1012-1037."""
    tmp = S - np.sum(IMF, axis=0)
    rng = np.max(tmp) - np.min(tmp)
    if ties is not None and _near(rng, RANGE_THR):
        ties.append(NearTie(where[0], where[1], "range", float(rng), RANGE_THR))
    if rng < RANGE_THR:
        return True
    tp = np.sum(np.abs(tmp))
    if ties is not None and _near(tp, TOTAL_POWER_THR):
        ties.append(NearTie(where[0], where[1], "total_power", float(tp), TOTAL_POWER_THR))
    if tp < TOTAL_POWER_THR:
        return True
    return False


# ---------------------------------------------------------------------------------------
# Whole decomposition
# ---------------------------------------------------------------------------------------
@dataclass
class EMDResult:
    imfs: np.ndarray                 # [n_imfs, N] (residue appended as last row if nonzero)
    n_imfs: int
    overflow: int
    debug_n: np.ndarray              # [max_imfs] sift count per IMF slot (-1 = not reached)
    debug_extno: np.ndarray
    debug_nzm: np.ndarray
    total_sifts: int = 0             # spline/update iterations performed
    peak_extrema: int = 0            # max over extrema passes of max(n_max, n_min)
    knot_sum: int = 0                # sum over sifts of (max knots + min knots)
    imf_cap_hit: int = 0
    near_ties: list = field(default_factory=list)


def emd(S, max_imf: int = -1, max_iteration: int = 1000, nbsym: int = 2,
        max_extrema: int | None = None, debug_slots: int = 16,
        mode: str = "pyemd", T=None) -> EMDResult:
    """PyEMD ``EMD.emd`` with the default stopping rule (FIXE = FIXE_H = 0).

    T = None: the time axis is arange(N), as PyEMD's 'simple' mode always makes it.  A
    given T is normalised (``normalize_time``) and used the way the reference kernel uses
    its time axis: knot positions, mirrored positions and the envelope evaluation points
    (emd.py:83-88; kernels.py:471-472, 528-552, 703); extrema are still found by sample
    index.  Reference counterpart:
    emd_batch_kernel (kernels.py:834-1080); debug records follow kernels.py:901-928,
    972-979 (sift count at convergence, = max_iteration with extno/nzm -1 when capped,
    -1 for slots never reached).  ``mode='reference'`` applies the reference kernel's own
    extrema, 3-knot and running-residue rules (module docstring).
    """
    if mode not in ("pyemd", "reference"):
        raise ValueError(f"mode must be 'pyemd' or 'reference', got {mode!r}")
    ref = mode == "reference"
    find = find_extrema_reference if ref else find_extrema_simple
    S = np.asarray(S, dtype=np.float64)
    N = S.size
    if T is None:
        T = np.arange(N, dtype=np.float64)
    else:
        if np.asarray(T).shape != (N,):
            raise ValueError(f"T has shape {np.asarray(T).shape}, the signal {S.shape}")
        T = normalize_time(T)
    dbg_n = np.full(debug_slots, -1, np.int64)
    dbg_e = np.full(debug_slots, -1, np.int64)
    dbg_z = np.full(debug_slots, -1, np.int64)
    ties: list = []
    if N == 0:
        # an empty trace has nothing to decompose (PyEMD cannot take one; the reference
        # kernel reports zero rows)
        return EMDResult(np.zeros((0, 0)), 0, 0, dbg_n, dbg_e, dbg_z)
    overflow = 0
    total_sifts = 0
    peak = 0
    knot_sum = 0

    def extrema(x, where=None):
        nonlocal overflow, peak
        imax, imin, nz = find(x)
        if where is not None and x.size:
            # sign tests at a rounding-level distance from zero: a sample (zero crossing)
            # or a first difference (extremum) within 1e-12 * max|x| of 0
            tol = NEAR_TIE_REL * float(np.max(np.abs(x)))
            nzs = np.count_nonzero(np.abs(x) <= tol)
            nzd = np.count_nonzero(np.abs(np.diff(x)) <= tol) if x.size > 1 else 0
            if nzs:
                ties.append(NearTie(where[0], where[1], "sample_sign", float(nzs), 0.0))
            if nzd:
                ties.append(NearTie(where[0], where[1], "difference_sign", float(nzd), 0.0))
        peak = max(peak, len(imax), len(imin))
        if max_extrema is not None:
            if len(imax) >= max_extrema or len(imin) >= max_extrema:
                overflow = 1
            imax, imin = imax[:max_extrema], imin[:max_extrema]
        return imax, imin, nz

    IMF = np.empty((0, N))
    imfNo = 0
    extNo = -1
    finished = False
    cap_hit = 0
    run_res = S.copy()  # reference mode: the running residue
    while not finished:
        residue = run_res if ref else S - np.sum(IMF[:imfNo], axis=0)
        imf = residue.copy()
        n = 0
        while True:
            n += 1
            if n >= max_iteration:
                overflow = 1
                if imfNo < debug_slots:
                    dbg_n[imfNo], dbg_e[imfNo], dbg_z[imfNo] = n, -1, -1
                break
            imax, imin, nzm = extrema(imf, None if (imfNo == 0 and n == 1) else (imfNo, n))
            extNo = len(imax) + len(imin)
            if extNo > 2 and len(imax) >= 1 and len(imin) >= 1:
                max_k, min_k = prepare_points_simple(imf, imax, imin, nbsym, T)
                max_env = spline_points(T, max_k, mode)
                min_env = spline_points(T, min_k, mode)
                total_sifts += 1
                knot_sum += max_k.shape[1] + min_k.shape[1]
                mean = 0.5 * (max_env + min_env)
                imf_old = imf.copy()
                imf = imf - mean
                imax2, imin2, nzm = extrema(imf, (imfNo, n))
                extNo = len(imax2) + len(imin2)
                f1 = check_imf(imf, imf_old, max_k, min_k, ties, (imfNo, n))
                f2 = abs(extNo - nzm) < 2
                if f1 and f2:
                    if imfNo < debug_slots:
                        dbg_n[imfNo], dbg_e[imfNo], dbg_z[imfNo] = n, extNo, nzm
                    break
            else:
                finished = True
                break
        IMF = np.vstack((IMF, imf.copy()))
        imfNo += 1
        if ref:
            run_res = run_res - imf
            rng = np.max(run_res) - np.min(run_res)
            if rng < RANGE_THR or np.sum(np.abs(run_res)) < TOTAL_POWER_THR:
                finished = True
                break
        elif end_condition(S, IMF, ties, (imfNo - 1, n)):
            finished = True
            break
        if imfNo == max_imf:
            cap_hit = 0 if finished else 1
            finished = True
            break
    if extNo <= 2 and IMF.shape[0] > 0:
        if ref:
            run_res = run_res + IMF[-1]  # the reference undoes the trend's subtraction
        IMF = IMF[:-1]
    residue = run_res if ref else S - np.sum(IMF, axis=0)
    if not np.allclose(residue, 0):
        IMF = np.vstack((IMF, residue))
    return EMDResult(IMF, IMF.shape[0], overflow, dbg_n, dbg_e, dbg_z, total_sifts, peak,
                     knot_sum, cap_hit, ties)


def envelopes(S, nbsym: int = 2, max_extrema: int | None = None, T=None):
    """One envelope pass without sifting (the reference's diagnostic
    ``envelope_batch_kernel``, kernels.py:1100-1166): extrema, mirror, both splines.
    T as in ``emd`` (None = arange(N); a given T is normalised).

    Returns dict(max_env, min_env, n_max, n_min, degenerate, max_knots, min_knots).
    """
    S = np.asarray(S, dtype=np.float64)
    N = S.size
    T = np.arange(N, dtype=np.float64) if T is None else normalize_time(T)
    imax, imin, _ = find_extrema_simple(S)
    if max_extrema is not None:
        imax, imin = imax[:max_extrema], imin[:max_extrema]
    out = dict(n_max=len(imax), n_min=len(imin), degenerate=0,
               max_env=np.zeros(N), min_env=np.zeros(N),
               max_knots=np.zeros((2, 0)), min_knots=np.zeros((2, 0)))
    if len(imax) + len(imin) <= 2 or len(imax) < 1 or len(imin) < 1:
        out["degenerate"] = 1
        return out
    mk, nk = prepare_points_simple(S, imax, imin, nbsym, T)
    out["max_env"] = spline_points(T, mk)
    out["min_env"] = spline_points(T, nk)
    out["max_knots"], out["min_knots"] = mk, nk
    return out


def emd_batch(signals, max_imfs: int = 16, max_iteration: int = 1000, nbsym: int = 2,
              max_extrema: int | None = None, T=None):
    """Batch driver returning the reference's padded outputs
    (imfs [B,max_imfs,N], n_imfs, overflow, debug_n/extno/nzm [B,max_imfs]) plus the
    per-signal EMDResult list.  The IMF cap is the reference's ``max_imfs - 1``
    (kernels.py:894) expressed as PyEMD ``max_imf``.  T: None, one time axis [N] shared
    by the batch, or one per signal [B,N] (each normalised on its own)."""
    signals = np.atleast_2d(np.asarray(signals, dtype=np.float64))
    B, N = signals.shape
    if T is not None:
        T = np.asarray(T, dtype=np.float64)
        if T.shape not in ((N,), (B, N)):
            raise ValueError(f"T has shape {T.shape}; expected ({N},) or ({B}, {N})")
    imfs = np.zeros((B, max_imfs, N))
    n_imfs = np.zeros(B, np.int32)
    ovf = np.zeros(B, np.int32)
    dn = np.full((B, max_imfs), -1, np.int32)
    de = np.full((B, max_imfs), -1, np.int32)
    dz = np.full((B, max_imfs), -1, np.int32)
    results = []
    for b in range(B):
        Tb = None if T is None else (T if T.ndim == 1 else T[b])
        r = emd(signals[b], max_imf=max_imfs - 1, max_iteration=max_iteration, nbsym=nbsym,
                max_extrema=max_extrema, debug_slots=max_imfs, T=Tb)
        k = r.n_imfs
        imfs[b, :k] = r.imfs
        n_imfs[b] = k
        ovf[b] = r.overflow
        dn[b], de[b], dz[b] = r.debug_n, r.debug_extno, r.debug_nzm
        results.append(r)
    return imfs, n_imfs, ovf, dn, de, dz, results
