"""Host-side checks of the boundary that need no GPU: the library exports, the T rule."""
import warnings

import numpy as np
import pytest

from cuda_emd.kernels import warn_nonuniform_T


def test_uniform_T_is_silent():
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        assert not warn_nonuniform_T(np.arange(100, dtype=np.float64))
        assert not warn_nonuniform_T(0.004 * np.arange(2001) + 3.0)
        assert not warn_nonuniform_T(np.arange(2, dtype=np.float64))


@pytest.mark.parametrize("T", [
    np.r_[0.0, 1.0, 2.0, 4.0, 5.0],             # one gap
    np.cumsum(np.linspace(1.0, 2.0, 50)),       # smoothly stretched
    np.arange(10, dtype=np.float64)[::-1],      # decreasing
])
def test_nonuniform_T_warns(T):
    with pytest.warns(UserWarning, match="not uniformly spaced"):
        assert warn_nonuniform_T(T)
