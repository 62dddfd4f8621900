"""Energy-level selection: eigenvalue intervals and filters.

Host-side logic of python/bipp/parameter_estimator.py:18-105 and python/bipp/filter.py:7-51.
The intervals decide which eigenpairs enter which image; the fused GPU pipeline applies
the same closed-interval test to the eigenvalues rounded to float32 (the precision the
reference stores them in, src/dataset_file.cpp:41-61).
"""
from __future__ import annotations

import numpy as np


def apply_filter(name: str, d):
    """python/bipp/filter.py:7-26."""
    d = np.array(d, copy=True)
    if name == "lsq":
        return d
    if name == "std":
        return np.sign(d)
    if name == "sqrt":
        sg = np.sign(d)
        return sg * np.sqrt(sg * d)
    if name == "inv_sqrt":
        sg = np.sign(d)
        d = sg * np.sqrt(sg * d)
        nz = d != 0
        d[nz] = 1.0 / d[nz]
        return d
    if name == "inv":
        nz = d != 0
        d[nz] = 1.0 / d[nz]
        return d
    raise ValueError(f"Unknown filter: {name}")


class Filter:
    """python/bipp/filter.py:29-51: closed-interval mask, then the named filter."""

    VALID = ("lsq", "std", "sqrt", "inv", "inv_sqrt")

    def __init__(self, filter_name: str, lower_bound: float, upper_bound: float) -> None:
        if filter_name not in self.VALID:
            raise ValueError(f"Unknown filter: {filter_name}")
        self.filter_name = filter_name
        self.lower_bound = lower_bound
        self.upper_bound = upper_bound

    def __call__(self, d):
        d = np.array(d, copy=True)
        d *= (d >= self.lower_bound) * (d <= self.upper_bound)
        return apply_filter(self.filter_name, d)


def centroid_to_intervals(centroid=None, d_min=0.0, d_max=float(np.finfo("f").max)):
    """Intervals between neighbouring centroids (python/bipp/parameter_estimator.py:18-52);
    row i covers the i-th largest centroid."""
    if centroid is None or np.asarray(centroid).size <= 1:
        return np.array([[d_min, d_max]])
    c = np.asarray(centroid, dtype=np.float64)
    order = np.argsort(c)
    srt = c[order]
    out = np.empty((c.size, 2))
    for i in range(c.size):
        k = order[i]
        out[i, 0] = d_min if k == 0 else 0.5 * (srt[k] + srt[k - 1])
        out[i, 1] = d_max if k == c.size - 1 else 0.5 * (srt[k] + srt[k + 1])
    return out


def infer_intervals(n_level: int, sigma: float, cluster_func: str, d_min: float, d_max: float,
                    d_all):
    """KMeans (random_state 42) on the eigenvalues, as python/bipp/parameter_estimator.py:55-105."""
    from sklearn.cluster import KMeans

    if cluster_func == "log" and d_min < 0:
        raise ValueError("log clustering requires a non-negative d_min")
    d = np.asarray(d_all, dtype=np.float64).ravel()
    d = d[d != 0.0]
    d = d[(d >= d_min) & (d <= d_max)]
    if 0 <= sigma < 1:
        n_remove = int((1 - sigma) * d.shape[0])
        if n_remove <= d.shape[0]:
            d = np.sort(d)[0:-n_remove]
    d = d.reshape(-1, 1)
    if cluster_func == "log":
        km = KMeans(n_clusters=n_level, random_state=42).fit(np.log(d))
        cent = np.sort(np.exp(km.cluster_centers_)[:, 0])[::-1]
    elif cluster_func == "none":
        km = KMeans(n_clusters=n_level, random_state=42).fit(d)
        cent = np.sort(km.cluster_centers_[:, 0])[::-1]
    else:
        raise ValueError(f'Unknown cluster function "{cluster_func}".')
    return centroid_to_intervals(cent, d_min, d_max)
