"""Level selection of the product (bipp_amd.levels) against the oracle's restatement of
python/bipp/filter.py and parameter_estimator.py, including the reference's row order."""
import numpy as np
import pytest

from bipp_amd import levels as L
from oracle import bipp_oracle as O


@pytest.mark.parametrize("n_level", [2, 3, 4, 6])
@pytest.mark.parametrize("sigma", [1.0, 0.95])
def test_intervals_equal_oracle(n_level, sigma):
    r = np.random.default_rng(n_level)
    d = np.abs(r.lognormal(0, 2, size=400))
    d[::13] = 0
    for func in ("log", "none"):
        assert np.array_equal(L.infer_intervals(n_level, sigma, func, 0.0, float("inf"), d),
                              O.infer_intervals(n_level, sigma, func, 0.0, float("inf"), d))


def test_centroid_rows_follow_reference_indexing():
    """Rows take the bounds at argsort(centroid)[i] of the ascending list: for descending
    centroids (what infer_intervals passes) row 0 is the largest centroid's interval."""
    c = np.array([100.0, 10.0, 1.0])
    iv = L.centroid_to_intervals(c, 0.0, 1e3)
    assert np.array_equal(iv, [[55.0, 1e3], [5.5, 55.0], [0.0, 5.5]])
    r = np.random.default_rng(3)
    for _ in range(10):
        c = r.normal(size=int(r.integers(2, 8)))
        assert np.array_equal(L.centroid_to_intervals(c, -9, 9), O.centroid_to_intervals(c, -9, 9))
    assert np.array_equal(L.centroid_to_intervals(None, 1, 2), [[1, 2]])


@pytest.mark.parametrize("name", ["lsq", "std", "sqrt", "inv", "inv_sqrt"])
def test_filters_equal_oracle(name):
    x = np.array([4.0, -9.0, 0.0, 2.5, 1e-3])
    assert np.allclose(L.apply_filter(name, x), O.apply_filter(name, x), rtol=1e-15)
    assert np.array_equal(L.Filter(name, -5, 3)(x), O.level_filter(name, -5, 3, x))


def test_unknown_names_rejected():
    with pytest.raises(ValueError):
        L.Filter("median", 0, 1)
    with pytest.raises(ValueError):
        L.infer_intervals(2, 1.0, "kmeans", 0.0, 1.0, [1.0, 2.0])
