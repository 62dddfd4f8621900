"""CPU oracle for the Bluebild synthesis hot path -- TEST INFRASTRUCTURE ONLY.

A numpy/scipy restatement of the reference's per-sample path, used as the checker by
`tests/`, by `__graft_entry__.smoke()` and by the CPU-baseline leg of `bench.py`.
Nothing in the product path imports this module.

Each function cites the reference code it restates.  How each one is pinned
(independently of retyping its formula) is listed in `PINNING` below and exercised by
tests/test_oracle.py:
  * flag_beams / eigh_gram bookkeeping: the LOFAR fixture counts (54 of 55 beams kept,
    nVis 2842/2840) and hand-built flag patterns;
  * gram_matrix: closed-form special cases (diagonal 1, sinc zeros at |d| = k lambda/2,
    W = I reduces to the sinc matrix, W-congruence) and brute force;
  * hegv: LAPACK ?hegv itself through scipy, plus residual/orthogonality invariants;
  * virtual_vis: Hermitian, trace identity, phase invariance of eigenvectors;
  * nuft_sum vs steering_image: two independent formulas of the same image agree,
    the uvw baseline convention is pinned by the fixture (uvw = 2pi/lambda (p_i - p_j));
  * infer_intervals: sklearn KMeans (the reference's own dependency) on tiny inputs
    with a hand-checkable answer.
"""
from __future__ import annotations

import math
from typing import Optional, Sequence, Tuple

import numpy as np
import scipy.linalg.lapack as lapack

PINNING = {
    "flag_beams": "LOFAR fixture counts + hand-built patterns",
    "gram_matrix": "closed-form special cases + brute force",
    "hegv": "LAPACK ?hegv via scipy + invariants",
    "eigh_gram": "composition of the above; LOFAR fixture counts",
    "eigh": "LAPACK ?heev via scipy + invariants",
    "virtual_vis": "Hermitian / trace / phase-invariance identities",
    "nuft_sum": "agreement with steering_image (independent formula)",
    "steering_image": "agreement with nuft_sum; positivity; source peak location",
    "infer_intervals": "sklearn KMeans on inputs with a known answer",
    "apply_filter": "closed forms",
}

TWO_PI = 2.0 * math.pi


def eps_of(dtype) -> float:
    return float(np.finfo(np.float32 if np.dtype(dtype) in (np.float32, np.complex64)
                          else np.float64).eps)


# --------------------------------------------------------------------------------------
# Stage A: flag -> Gram -> ?hegv -> un-beam          (src/eigensolver.cpp:129-225)
# --------------------------------------------------------------------------------------
def flag_beams(s: np.ndarray, real_dtype=np.float32, mode: str = "gram"):
    """Live-entry test on the lower triangle (src/eigensolver.cpp:148-159).

    mode "gram": |s_rc|^2 > eps (eigh_gram, :153); mode "eigh": |Re|>=eps or |Im|>=eps
    (eigh, :62-63).  Returns (kept beam indices, nVis)."""
    eps = eps_of(real_dtype)
    rdt = np.dtype(real_dtype).type
    nb = s.shape[0]
    low = np.tril(np.ones((nb, nb), dtype=bool))
    if mode == "gram":
        # std::norm: re^2 + im^2 in the working precision, no square root (:153)
        re = s.real.astype(rdt)
        im = s.imag.astype(rdt)
        live = (re * re + im * im > rdt(eps)) & low
    else:
        live = ((np.abs(s.real) >= eps) | (np.abs(s.imag) >= eps)) & low
    flag = live.any(axis=0) | live.any(axis=1)
    n_vis = int(live.sum() * 2 - np.trace(live))
    return np.nonzero(flag)[0], n_vis


def sinc_matrix(xyz: np.ndarray, wl: float) -> np.ndarray:
    """sinc(2|p_i - p_j| / lambda), sinc(x) = sin(pi x)/(pi x), diagonal 1
    (src/host/gram_matrix.cpp:23-27,47-56)."""
    d = np.sqrt(((xyz[:, None, :] - xyz[None, :, :]) ** 2).sum(-1))
    return np.sinc(2.0 * d / wl)


def gram_matrix(w: np.ndarray, xyz: np.ndarray, wl: float, prefactor: float = 1.0) -> np.ndarray:
    """G = W^H Sigma W (src/host/gram_matrix.cpp:29-64).  The C++ reference has no 4 pi
    prefactor; python/bipp/gram.py:110 has one -- `prefactor` selects."""
    sig = prefactor * sinc_matrix(xyz, wl)
    return w.conj().T @ (sig @ w)


def hegv(a: np.ndarray, b: np.ndarray, single: bool):
    """LAPACK ?hegv(itype=1, jobz='V', uplo='L') (src/host/lapack_api.hpp:112-158).
    Returns (ascending eigenvalues, eigenvectors with Z^H B Z = I)."""
    if single:
        w, z, info = lapack.chegv(a.astype(np.complex64), b.astype(np.complex64), itype=1,
                                  jobz="V", uplo="L")
    else:
        w, z, info = lapack.zhegv(a.astype(np.complex128), b.astype(np.complex128), itype=1,
                                  jobz="V", uplo="L")
    if info != 0:
        raise np.linalg.LinAlgError(f"?hegv info={info}")
    return w, z


def heev(a: np.ndarray, single: bool):
    """LAPACK ?heev(jobz='V', uplo='L') (src/host/lapack_api.hpp:70-110)."""
    if single:
        w, z, info = lapack.cheev(a.astype(np.complex64), compute_v=1, lower=1)
    else:
        w, z, info = lapack.zheev(a.astype(np.complex128), compute_v=1, lower=1)
    if info != 0:
        raise np.linalg.LinAlgError(f"?heev info={info}")
    return w, z


def _lower_to_hermitian(m: np.ndarray) -> np.ndarray:
    low = np.tril(m)
    return low + np.tril(m, -1).conj().T


def eigh_gram(wl: float, s: np.ndarray, w: Optional[np.ndarray], xyz: np.ndarray,
              single: bool = False, prefactor: float = 1.0):
    """bipp::eigh_gram (src/eigensolver.cpp:129-225).

    Returns (v (nAntenna, nBeam), d (nBeam,), nBeamReduced, scale).  Only the lower
    triangle of s is read.  d is ascending over the kept beams and zero-padded."""
    nb = s.shape[0]
    n_ant = xyz.shape[0]
    if w is None:
        w = np.eye(n_ant, nb)
    rdt = np.float32 if single else np.float64
    cdt = np.complex64 if single else np.complex128
    s = s.astype(cdt)
    w = w.astype(cdt)
    idx, n_vis = flag_beams(s, rdt, "gram")
    nr = idx.size
    d = np.zeros(nb, dtype=rdt)
    v = np.zeros((n_ant, nb), dtype=cdt)
    if nr:
        s_red = _lower_to_hermitian(s[np.ix_(idx, idx)])
        w_red = w[:, idx]
        g = gram_matrix(w_red, xyz.astype(rdt), rdt(wl), prefactor)
        ev, z = hegv(s_red, g, single)
        d[:nr] = ev
        v[:, :nr] = w_red @ z.astype(cdt)
    scale = (1.0 / n_vis) if n_vis else 0.0
    return v, d, nr, rdt(scale)


def eigh(wl: float, s: np.ndarray, w: Optional[np.ndarray], single: bool = False,
         n_antenna: Optional[int] = None):
    """bipp::eigh (src/eigensolver.cpp:40-127): standard problem, different flag test."""
    nb = s.shape[0]
    if w is None:
        w = np.eye(n_antenna or nb, nb)
    n_ant = w.shape[0]
    rdt = np.float32 if single else np.float64
    cdt = np.complex64 if single else np.complex128
    idx, n_vis = flag_beams(s.astype(cdt), rdt, "eigh")
    nr = idx.size
    d = np.zeros(nb, dtype=rdt)
    v = np.zeros((n_ant, nb), dtype=cdt)
    if nr:
        s_red = _lower_to_hermitian(s.astype(cdt)[np.ix_(idx, idx)])
        ev, z = heev(s_red, single)
        d[:nr] = ev
        v[:, :nr] = w.astype(cdt)[:, idx] @ z.astype(cdt)
    scale = (1.0 / n_vis) if n_vis else 0.0
    return v, d, nr, rdt(scale)


# --------------------------------------------------------------------------------------
# Levels and filters  (python/bipp/filter.py, python/bipp/parameter_estimator.py)
# --------------------------------------------------------------------------------------
def apply_filter(name: str, d: np.ndarray) -> np.ndarray:
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
    raise ValueError(name)


def level_filter(name: str, lo: float, hi: float, d: np.ndarray) -> np.ndarray:
    """Filter.__call__ (python/bipp/filter.py:48-51): closed interval mask, then f."""
    d = np.array(d, copy=True)
    d *= (d >= lo) * (d <= hi)
    return apply_filter(name, d)


def centroid_to_intervals(centroid=None, d_min=0.0, d_max=float(np.finfo("f").max)):
    """python/bipp/parameter_estimator.py:18-52 (including its index arithmetic)."""
    if centroid is None or centroid.size <= 1:
        return np.array([[d_min, d_max]])
    intervals = np.empty((centroid.size, 2))
    sorted_idx = np.argsort(centroid)
    sorted_centroid = centroid[sorted_idx]
    for i in range(centroid.size):
        idx = sorted_idx[i]
        intervals[i, 0] = d_min if idx == 0 else (sorted_centroid[idx] + sorted_centroid[idx - 1]) / 2
        intervals[i, 1] = (d_max if idx == centroid.size - 1
                           else (sorted_centroid[idx] + sorted_centroid[idx + 1]) / 2)
    return intervals


def infer_intervals(n_level: int, sigma: float, cluster_func: str, d_min: float, d_max: float,
                    d_all) -> np.ndarray:
    """python/bipp/parameter_estimator.py:55-105 (sklearn KMeans, random_state=42)."""
    import sklearn.cluster as skcl

    if cluster_func == "log" and d_min < 0:
        raise ValueError("log clustering needs d_min >= 0")
    d_all = np.array(d_all).flatten()
    d_all = d_all[d_all != 0.0]
    d_all = d_all[d_all >= d_min]
    d_all = d_all[d_all <= d_max]
    if 0 <= sigma < 1:
        n_remove = int((1 - sigma) * d_all.shape[0])
        if n_remove <= d_all.shape[0]:
            d_all = np.sort(d_all)
            d_all = d_all[0:-n_remove]
    d_all = d_all.reshape(-1, 1)
    if cluster_func == "log":
        km = skcl.KMeans(n_clusters=n_level, random_state=42).fit(np.log(d_all))
        cent = np.sort(np.exp(km.cluster_centers_)[:, 0])[::-1]
    elif cluster_func == "none":
        km = skcl.KMeans(n_clusters=n_level, random_state=42).fit(d_all)
        cent = np.sort(km.cluster_centers_[:, 0])[::-1]
    else:
        raise ValueError(cluster_func)
    return centroid_to_intervals(cent, d_min, d_max)


# --------------------------------------------------------------------------------------
# Stage B: virtual visibilities and images
# --------------------------------------------------------------------------------------
def virtual_vis(scale: float, d_masked: np.ndarray, v: np.ndarray) -> np.ndarray:
    """V_E diag(scale d_E) V_E^H over nonzero d, flattened column-major so that entry
    (i, j) sits at i + N j (src/host/virtual_vis.cpp:19-73)."""
    nz = d_masked != 0
    ve = v[:, nz]
    vv = (ve * (scale * d_masked[nz])) @ ve.conj().T
    return vv.reshape(-1, order="F")


def baselines_uvw(xyz: np.ndarray) -> np.ndarray:
    """(N^2, 3) baselines with uvw[i + N j] = p_i - p_j (python/bipp/instrument.py:246-251,
    python/bipp/frame.py:75-89)."""
    b = xyz[:, None, :] - xyz[None, :, :]          # b[i, j] = p_i - p_j
    return b.transpose(1, 0, 2).reshape(-1, 3)      # row j*N + i


def nuft_sum(alpha: float, values: np.ndarray, uvw: np.ndarray, lmn: np.ndarray,
             chunk: int = 4096) -> np.ndarray:
    """out[q] = sum_j Re(c_j exp(i alpha (x_q u_j + y_q v_j + z_q w_j)))
    (src/host/kernels/nuft_sum.cpp:9-31), evaluated in float64."""
    out = np.zeros(lmn.shape[0])
    for q0 in range(0, lmn.shape[0], chunk):
        ph = alpha * (lmn[q0:q0 + chunk] @ uvw.T)
        out[q0:q0 + chunk] = (np.cos(ph) * values.real - np.sin(ph) * values.imag).sum(1)
    return out


def steering_image(v: np.ndarray, weights: np.ndarray, xyz: np.ndarray, wl: float,
                   lmn: np.ndarray, chunk: int = 4096) -> np.ndarray:
    """I(x) = sum_e w_e |sum_i V_ie exp(+i 2pi/lambda p_i.x)|^2 -- the exact identity
    sum_ij VV_ij e^{i uvw_ij.x} with VV = V diag(w) V^H and uvw_ij = 2pi/lambda (p_i - p_j)
    (src/host/kernels/nuft_sum.cpp:20-26 with src/host/virtual_vis.cpp:52-68).
    weights: (..., K) -> output (..., P)."""
    w2 = np.atleast_2d(weights)
    out = np.zeros((w2.shape[0], lmn.shape[0]))
    k = TWO_PI / wl
    for q0 in range(0, lmn.shape[0], chunk):
        e = np.exp(1j * k * (xyz @ lmn[q0:q0 + chunk].T))       # (N, Pc)
        b = v.T @ e                                              # (K, Pc)
        out[:, q0:q0 + chunk] = w2 @ (np.abs(b) ** 2)
    return out.reshape(weights.shape[:-1] + (lmn.shape[0],))


def sample_level_weights(d: np.ndarray, scale: float, intervals: np.ndarray,
                         filt: str = "lsq", d_storage=np.float32) -> np.ndarray:
    """Per-level weights scale * f(d [lo <= d <= hi]) of one sample, computed on the
    stored-precision eigenvalues as the reference does (dataset eigenvalues are
    float32: src/dataset_file.cpp:41-61; mask: tests/test_nufft_synthesis_lofar.cpp:112-125,
    python/bipp/filter.py:48-51)."""
    ds = np.asarray(d, dtype=d_storage)
    out = np.zeros((intervals.shape[0], d.shape[0]))
    for lvl, (lo, hi) in enumerate(intervals):
        m = (ds >= d_storage(lo)) & (ds <= d_storage(hi))
        out[lvl] = scale * apply_filter(filt, np.where(m, np.asarray(d, np.float64), 0.0))
    return out


def level_masks(d: np.ndarray, intervals: np.ndarray, d_storage=np.float32) -> np.ndarray:
    ds = np.asarray(d, dtype=d_storage)
    return np.stack([(ds >= d_storage(lo)) & (ds <= d_storage(hi)) & (ds != 0)
                     for lo, hi in intervals])


def synthesize(samples, intervals: np.ndarray, lmn: np.ndarray, filt: str = "lsq",
               single_eig: bool = False, normalize: bool = True, method: str = "steering"):
    """Full per-sample chain for a list of synthetic samples (W = I):
    eigh_gram -> level weights -> image (steering identity or direct NUFT sum of the
    virtual visibilities) -> sum over samples -> 1/nSamples
    (src/nufft_synthesis.cpp:134-165, src/image_synthesis.cpp:152-161).
    Returns (images (L, P), per-sample eigenvalues list)."""
    imgs = np.zeros((intervals.shape[0], lmn.shape[0]))
    evs = []
    for smp in samples:
        v, d, nr, scale = eigh_gram(smp.wl, smp.s, None, smp.xyz, single=single_eig)
        evs.append(d)
        wts = sample_level_weights(d, float(scale), intervals, filt)
        if method == "steering":
            imgs += steering_image(v.astype(np.complex128), wts, smp.xyz, smp.wl, lmn)
        else:
            uvw = baselines_uvw(smp.xyz) * (TWO_PI / smp.wl)
            for lvl in range(intervals.shape[0]):
                vv = virtual_vis(1.0, wts[lvl], v.astype(np.complex128))
                imgs[lvl] += nuft_sum(1.0, vv, uvw, lmn)
    if normalize and samples:
        imgs /= len(samples)
    return imgs, evs


def rel_l2(a: np.ndarray, ref: np.ndarray) -> float:
    den = np.linalg.norm(ref)
    return float(np.linalg.norm(np.asarray(a, np.float64) - ref) / (den if den else 1.0))
