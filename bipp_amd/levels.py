"""Energy levels: which eigenpairs feed which image, and with what weight.

A level is a closed interval [lo, hi] on the eigenvalues; its image uses f(d) for the
eigenvalues d inside the interval and nothing else.  f is one of five filters: the
eigenvalue itself ("lsq"), its sign ("std"), its signed square root ("sqrt") and their
reciprocals ("inv", "inv_sqrt"); zero stays zero.  Intervals are placed between the centres
of a one-dimensional k-means clustering of the (log) eigenvalues.

Restates python/bipp/filter.py:7-51 and python/bipp/parameter_estimator.py:18-105 -- the
same masks, filters and interval placement, including the reference's ordering of the
interval rows.  The fused GPU pipeline applies the same closed-interval test to the
eigenvalues rounded to float32, the precision a dataset stores them in
(src/dataset_file.cpp:41-61).
"""
from __future__ import annotations

from typing import Callable, Dict

import numpy as np


def _signed_root(x: np.ndarray) -> np.ndarray:
    return np.sign(x) * np.sqrt(np.abs(x))


def _reciprocal(x: np.ndarray) -> np.ndarray:
    out = np.zeros_like(x)
    nz = x != 0
    out[nz] = 1.0 / x[nz]
    return out


_FILTERS: Dict[str, Callable[[np.ndarray], np.ndarray]] = {
    "lsq": lambda x: x,
    "std": np.sign,
    "sqrt": _signed_root,
    "inv": _reciprocal,
    "inv_sqrt": lambda x: _reciprocal(_signed_root(x)),
}


def apply_filter(name: str, d):
    """f(d) elementwise for the filter `name` (python/bipp/filter.py:7-26)."""
    try:
        fn = _FILTERS[name]
    except KeyError:
        raise ValueError(f"filter {name!r} is not one of {sorted(_FILTERS)}") from None
    return fn(np.array(d, dtype=np.result_type(np.asarray(d), np.float32), copy=True))


class Filter:
    """Callable level: f applied to the eigenvalues inside [lower_bound, upper_bound], zero
    elsewhere (python/bipp/filter.py:29-51)."""

    VALID = tuple(_FILTERS)

    def __init__(self, filter_name: str, lower_bound: float, upper_bound: float) -> None:
        if filter_name not in _FILTERS:
            raise ValueError(f"filter {filter_name!r} is not one of {sorted(_FILTERS)}")
        self.filter_name = filter_name
        self.lower_bound = lower_bound
        self.upper_bound = upper_bound

    def __call__(self, d):
        vals = np.array(d, copy=True)
        inside = (vals >= self.lower_bound) & (vals <= self.upper_bound)
        return apply_filter(self.filter_name, np.where(inside, vals, 0))


def centroid_to_intervals(centroid=None, d_min=0.0, d_max=float(np.finfo("f").max)):
    """Interval rows between neighbouring centroids (python/bipp/parameter_estimator.py:18-52).

    Boundaries sit halfway between consecutive centroids in ascending order, the outermost
    ones at d_min and d_max.  Row i takes the bounds found at position argsort(centroid)[i]
    of the ascending list -- the reference indexes its ascending centroids by that index,
    which for the descending centroids infer_intervals passes gives row i to the i-th
    largest centroid."""
    if centroid is None or np.asarray(centroid).size <= 1:
        return np.array([[d_min, d_max]])
    c = np.asarray(centroid, dtype=np.float64).ravel()
    order = np.argsort(c)
    asc = c[order]
    mids = (asc[1:] + asc[:-1]) / 2
    lower = np.concatenate(([d_min], mids))
    upper = np.concatenate((mids, [d_max]))
    return np.stack([lower[order], upper[order]], axis=1)


def infer_intervals(n_level: int, sigma: float, cluster_func: str, d_min: float, d_max: float,
                    d_all):
    """Level intervals from the eigenvalues of many samples
    (python/bipp/parameter_estimator.py:55-105): zeros and values outside [d_min, d_max]
    dropped; for 0 <= sigma < 1 the int((1 - sigma) n) largest values are dropped too (an
    empty slice when that count is zero, as in the reference); k-means with random_state 42
    on log(d) ("log") or d ("none"); centroids passed largest first."""
    from sklearn.cluster import KMeans

    if cluster_func not in ("log", "none"):
        raise ValueError(f"cluster function {cluster_func!r} is not 'log' or 'none'")
    if cluster_func == "log" and d_min < 0:
        raise ValueError("log clustering needs d_min >= 0")
    d = np.asarray(d_all, dtype=np.float64).ravel()
    d = d[(d != 0.0) & (d >= d_min) & (d <= d_max)]
    if 0 <= sigma < 1:
        drop = int((1 - sigma) * d.size)
        if drop <= d.size:
            d = np.sort(d)[:-drop] if drop else np.sort(d)[0:-0]
    x = d.reshape(-1, 1)
    if cluster_func == "log":
        x = np.log(x)
    centres = KMeans(n_clusters=n_level, random_state=42).fit(x).cluster_centers_[:, 0]
    if cluster_func == "log":
        centres = np.exp(centres)
    return centroid_to_intervals(np.sort(centres)[::-1], d_min, d_max)
