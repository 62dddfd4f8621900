"""ctypes access to oracle/libbipp_cpu.so -- TEST INFRASTRUCTURE ONLY (checker and CPU
baseline).  Build with `make -C oracle`."""
from __future__ import annotations

import ctypes as C
import glob
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbipp_cpu.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def openblas_path() -> str:
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(scipy.__file__), "..", "scipy.libs",
                                  "libscipy_openblas*.so"))
    if not libs:
        raise RuntimeError("scipy's bundled OpenBLAS not found")
    return os.path.abspath(libs[0])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, ip, fp, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.bippc_init_lapack.argtypes = [C.c_char_p]
        L.bippc_lapack_threads.argtypes = [C.c_int]
        L.bippc_max_threads.restype = C.c_int
        L.bippc_eigh_gram_c.argtypes = [C.c_float, C.c_int, vp, vp, vp, vp, ip, fp]
        L.bippc_eigh_gram_c.restype = C.c_int
        L.bippc_eigh_gram_batch.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, vp, vp, vp]
        L.bippc_eigh_gram_batch.restype = C.c_int
        L.bippc_steering_image.argtypes = [C.c_int, C.c_int, vp, vp, C.c_int, vp, C.c_int, vp, vp]
        L.bippc_nuft_sum_d.argtypes = [C.c_double, C.c_int, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        if L.bippc_init_lapack(openblas_path().encode()) != 0:
            raise RuntimeError("could not load LAPACK chegv")
        _lib = L
    return _lib


def eigh_gram_c(wl: float, s: np.ndarray, xyz: np.ndarray):
    """float32 eigh_gram with W = I; s (N, N) standard orientation, xyz (N, 3)."""
    n = s.shape[0]
    s_cm = np.asfortranarray(s, dtype=np.complex64)
    x_cm = np.asfortranarray(xyz, dtype=np.float32)
    d = np.zeros(n, np.float32)
    v = np.zeros((n, n), np.complex64, order="F")
    nr = C.c_int()
    sc = C.c_float()
    info = lib().bippc_eigh_gram_c(np.float32(wl), n, s_cm.ctypes.data, x_cm.ctypes.data,
                                   d.ctypes.data, v.ctypes.data, C.byref(nr), C.byref(sc))
    if info != 0:
        raise np.linalg.LinAlgError(f"chegv info={info}")
    return v, d, nr.value, np.float32(sc.value)


def steering_image(v: np.ndarray, weights: np.ndarray, pos_wl: np.ndarray, lmn: np.ndarray,
                   img: np.ndarray) -> None:
    """img (L, P) float32 += sum_e w[l, e] |sum_i v_ie exp(2 pi i pos_i . x)|^2."""
    v = np.asfortranarray(v, dtype=np.complex64)
    w = np.ascontiguousarray(weights, dtype=np.float32)
    pos = np.asfortranarray(pos_wl, dtype=np.float64)
    lm = np.asfortranarray(lmn, dtype=np.float32)
    assert img.dtype == np.float32 and img.flags.c_contiguous
    lib().bippc_steering_image(v.shape[0], v.shape[1], v.ctypes.data, w.ctypes.data, w.shape[0],
                               pos.ctypes.data, lm.shape[0], lm.ctypes.data, img.ctypes.data)


def nuft_sum_d(alpha: float, values: np.ndarray, uvw: np.ndarray, lmn: np.ndarray,
               out: np.ndarray) -> None:
    """out (float64, P) += sum_j Re(c_j exp(i alpha (x u + y v + z w))), OpenMP over P."""
    c = np.ascontiguousarray(values, dtype=np.complex128)
    u, v, w = (np.ascontiguousarray(uvw[:, i], dtype=np.float64) for i in range(3))
    x, y, z = (np.ascontiguousarray(lmn[:, i], dtype=np.float64) for i in range(3))
    assert out.dtype == np.float64 and out.flags.c_contiguous
    lib().bippc_nuft_sum_d(float(alpha), c.size, c.ctypes.data, u.ctypes.data, v.ctypes.data,
                           w.ctypes.data, x.size, x.ctypes.data, y.ctypes.data, z.ctypes.data,
                           out.ctypes.data)


def eigh_gram_batch(wl, s: np.ndarray, xyz: np.ndarray):
    """float32 eigh_gram of independent samples, OpenMP over samples with single-threaded
    LAPACK each; s (nS, N, N) standard orientation, xyz (nS, N, 3)."""
    ns, n = s.shape[0], s.shape[1]
    s_cm = np.ascontiguousarray(np.swapaxes(s, 1, 2), dtype=np.complex64)
    x_cm = np.ascontiguousarray(np.swapaxes(xyz, 1, 2), dtype=np.float32)
    w = np.ascontiguousarray(np.broadcast_to(np.asarray(wl, np.float32), (ns,)))
    d = np.zeros((ns, n), np.float32)
    v = np.zeros((ns, n, n), np.complex64)   # per sample column-major: v[k].T is (N, N)
    nr = np.zeros(ns, np.int32)
    sc = np.zeros(ns, np.float32)
    info = lib().bippc_eigh_gram_batch(ns, w.ctypes.data, n, s_cm.ctypes.data, x_cm.ctypes.data,
                                       d.ctypes.data, v.ctypes.data, nr.ctypes.data, sc.ctypes.data)
    if info != 0:
        raise np.linalg.LinAlgError(f"chegv info={info}")
    return np.swapaxes(v, 1, 2), d, nr, sc
