"""Host-side logic of the cost model's planner (no GPU): the held-out rank correlation that
decides whether predicted sifts enter the queue order."""
import types

import numpy as np
from scipy.stats import spearmanr

from cuda_emd import costmodel


def test_rank_corr_matches_spearman_with_ties():
    rng = np.random.default_rng(0)
    for n in (2, 7, 200):
        a = rng.integers(0, 5, n).astype(float)  # many ties
        b = a + rng.normal(0, 2, n)
        want = spearmanr(a, b).correlation
        got = costmodel.rank_corr(a, b)
        if np.isnan(want):
            assert np.isnan(got)
        else:
            assert abs(got - want) < 1e-12


def test_rank_corr_edge_cases():
    assert np.isnan(costmodel.rank_corr([1.0], [2.0]))
    assert np.isnan(costmodel.rank_corr([1.0, 1.0, 1.0], [1.0, 2.0, 3.0]))
    assert costmodel.rank_corr([1, 2, 3], [10, 20, 30]) == 1.0
    assert costmodel.rank_corr([1, 2, 3], [30, 20, 10]) == -1.0


def test_orders_by_sifts_threshold():
    prop = costmodel.CostModel.orders_by_sifts.fget
    assert prop(types.SimpleNamespace(sift_rank=float("nan")))  # no held-out evidence
    assert not prop(types.SimpleNamespace(sift_rank=costmodel.RANK_MIN - 0.01))
    assert prop(types.SimpleNamespace(sift_rank=costmodel.RANK_MIN))
    assert prop(types.SimpleNamespace(sift_rank=0.9))
