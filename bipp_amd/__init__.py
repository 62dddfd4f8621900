"""bipp_amd -- MI355X-native Bluebild synthesis with the interface of the `bipp` module.

Mirrors the pybind module of the reference (python/pybipp.cpp:285-493) for the hot path:
`Context`, `communicator`, `Partition`, `NufftSynthesisOptions`, `eigh_gram`, `eigh`,
`image_synthesis` on in-memory `Dataset` / `ImageProp` / `ImageData` objects (the HDF5 file
classes are not built), plus the fused `NufftSynthesis` collect/get pipeline.  All
computation runs in libbipp_amd.so on the GPU through the C-ABI of include/bipp_hip.h.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import sys
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi
from ._capi import (GenericError, call, check, lib, ptr, COMPUTE, FILTERS, IMAGERS,  # noqa: F401
                    PRECISION, PU)
from .levels import Filter, apply_filter, centroid_to_intervals, infer_intervals  # noqa: F401
from .dataset import (Dataset, ImageData, ImageProp, InMemoryDataset,  # noqa: F401
                      InMemoryImageData, InMemoryImageProp)

__version__ = "1.0.0-mi355x"

for _name in _capi.ERROR_NAMES.values():
    if hasattr(_capi, _name):
        globals()[_name] = getattr(_capi, _name)


@dataclasses.dataclass(frozen=True)
class CompileConfig:
    """python/pybipp.cpp:198-224 -- what the library was built with."""

    cuda: bool = False
    rocm: bool = True
    umpire: bool = False
    omp: bool = False
    mpi: bool = False


config = CompileConfig()


def device_count() -> int:
    c = C.c_int(0)
    check(lib().bipp_device_count(C.byref(c)))
    return c.value


class communicator:
    """Single-process stand-in for bipp::Communicator (include/bipp/communicator.hpp):
    this build runs on one GPU, so world() and local() both have one rank."""

    def __init__(self) -> None:
        self.size = 1
        self.rank = 0

    @staticmethod
    def world() -> "communicator":
        return communicator()

    @staticmethod
    def local() -> "communicator":
        return communicator()


class Context:
    """bipp.Context(pu) (python/pybipp.cpp:314-319): owns one HIP stream on the GPU."""

    def __init__(self, pu: str = "AUTO", comm: Optional[communicator] = None) -> None:
        key = pu.upper()
        if key not in PU:
            raise _capi._EXC[3](f"unknown processing unit {pu!r}")
        h = C.c_void_p()
        check(lib().bipp_ctx_create(PU[key], C.byref(h)))
        self._h = h
        self.comm = comm or communicator()

    @property
    def processing_unit(self) -> str:
        v = C.c_int()
        check(lib().bipp_ctx_processing_unit(self._h, C.byref(v)))
        return {0: "AUTO", 1: "CPU", 2: "GPU"}[v.value]

    @property
    def handle(self) -> C.c_void_p:
        return self._h

    def stream(self) -> int:
        s = C.c_void_p()
        check(lib().bipp_ctx_stream(self._h, C.byref(s)))
        return s.value or 0

    def synchronize(self) -> None:
        check(lib().bipp_ctx_synchronize(self._h))

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and not sys.is_finalizing():
            lib().bipp_ctx_destroy(h)
            self._h = None


class Partition:
    """bipp.Partition (include/bipp/image_synthesis.hpp:23-53)."""

    def __init__(self, method: str = "auto", dimensions=(1, 1, 1)) -> None:
        self.method = method
        self.dimensions = tuple(int(x) for x in dimensions)

    @staticmethod
    def auto() -> "Partition":
        return Partition("auto")

    @staticmethod
    def none() -> "Partition":
        return Partition("none")

    @staticmethod
    def grid(dimensions) -> "Partition":
        return Partition("grid", dimensions)

    def _c_args(self):
        if self.method not in _capi.PARTITION:
            raise _capi._EXC[3](f"unknown partition method {self.method!r}")
        dims = (C.c_size_t * 3)(*self.dimensions)
        return _capi.PARTITION[self.method], dims


class NufftSynthesisOptions:
    """bipp.NufftSynthesisOptions (include/bipp/image_synthesis.hpp:55-148)."""

    def __init__(self) -> None:
        self.tolerance = 0.001
        self.sample_batch_size: Optional[int] = None
        self.local_uvw_partition = Partition.auto()
        self.normalize_image = True
        self.apply_scaling = True
        self._precision = "single"
        # additions of this build: the compute form ("mixed" | "fp32" | "fp64"; default by
        # precision) and the image_synthesis imager ("nufft" | "direct")
        self.compute: Optional[str] = None
        self.imager = "nufft"

    @property
    def precision(self) -> str:
        return self._precision

    def set_precision(self, prec: str) -> "NufftSynthesisOptions":
        p = prec.lower()
        if p not in PRECISION:
            raise _capi._EXC[3](f"unknown precision {prec!r}")
        self._precision = p
        return self

    def set_tolerance(self, tol: float) -> "NufftSynthesisOptions":
        self.tolerance = float(tol)
        return self

    def set_sample_batch_size(self, size: Optional[int]) -> "NufftSynthesisOptions":
        self.sample_batch_size = size
        return self

    def set_local_uvw_partition(self, p: Partition) -> "NufftSynthesisOptions":
        self.local_uvw_partition = p
        return self

    def set_normalize_image(self, normalize: bool) -> "NufftSynthesisOptions":
        self.normalize_image = bool(normalize)
        return self

    def set_apply_scaling(self, apply: bool) -> "NufftSynthesisOptions":
        self.apply_scaling = bool(apply)
        return self

    def compute_mode(self) -> str:
        if self.compute:
            return self.compute
        return "fp64" if self._precision == "double" else "mixed"


def _fortran(a, dtype):
    return np.asfortranarray(np.asarray(a, dtype=dtype))


def _eigh_common(gram: bool, wl, s, w, xyz=None):
    s = np.asarray(s)
    single = s.dtype == np.complex64 or (s.dtype not in (np.complex128,) and
                                         isinstance(wl, np.float32))
    rdt, cdt = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    s = _fortran(s, cdt)
    if s.ndim != 2 or s.shape[0] != s.shape[1]:
        raise _capi._EXC[3]("s must be a square 2D array")
    n_beam = s.shape[0]
    if w is None:
        raise _capi._EXC[3]("w is required")
    w = _fortran(w, cdt)
    if w.ndim != 2 or w.shape[1] != n_beam:
        raise _capi._EXC[3]("w must be (nAntenna, nBeam)")
    n_ant = w.shape[0]
    d = np.zeros(n_beam, dtype=rdt)
    v = np.zeros((n_ant, n_beam), dtype=cdt, order="F")
    nr = C.c_size_t()
    sc = (C.c_float if single else C.c_double)()
    rp = C.POINTER(C.c_float if single else C.c_double)
    if gram:
        xyz = _fortran(xyz, rdt)
        if xyz.shape != (n_ant, 3):
            raise _capi._EXC[3]("xyz must be (nAntenna, 3)")
        fn = lib().bipp_eigh_gram_s if single else lib().bipp_eigh_gram_d
        check(fn(rdt(wl), n_ant, n_beam, ptr(s), n_beam, ptr(w), n_ant,
                 xyz.ctypes.data_as(rp), n_ant, d.ctypes.data_as(rp), ptr(v), n_ant,
                 C.byref(nr), C.byref(sc)))
    else:
        fn = lib().bipp_eigh_s if single else lib().bipp_eigh_d
        check(fn(rdt(wl), n_ant, n_beam, ptr(s), n_beam, ptr(w), n_ant,
                 d.ctypes.data_as(rp), ptr(v), n_ant, C.byref(nr), C.byref(sc)))
    return v, d, rdt(sc.value)


def eigh_gram(wl, s, w, xyz):
    """bipp.eigh_gram(wl, s, w, xyz) -> (v, d, scale) (python/pybipp.cpp:254-275,462-477).
    complex64 inputs select the single-precision solve, complex128 the double one."""
    return _eigh_common(True, wl, s, w, xyz)


def eigh(wl, s, w):
    """bipp.eigh(wl, s, w) -> (v, d, scale) (python/pybipp.cpp:234-252,479-493)."""
    return _eigh_common(False, wl, s, w)


class NufftSynthesis:
    """Fused per-sample Bluebild chain on the GPU: collect(samples) / get(level).

    Stage A (flag, Gram, generalized eigensolve) and stage B (level split, image) of
    src/eigensolver.cpp:129-225 and src/nufft_synthesis.cpp:74-170 run back to back on
    device for every collected sample; images accumulate on the GPU.

    compute: "mixed" (float64 eigensolve and phases, split-bf16 MFMA contraction),
             "fp32" (plain float32 form), "fp64" (float64 everywhere).
    imager:  "steering" (eigenvector x exp(i k p.x) contraction) or "direct"
             (virtual visibilities + direct NUFT sum, the reference's arithmetic).
    batched: one launch sequence per collect (True) or per sample (False).
    """

    def __init__(self, ctx: Context, n_antenna: int, n_beam: int, intervals, lmn,
                 filter_name: str = "lsq", compute: str = "mixed", imager: str = "steering",
                 batched: bool = True, tolerance: Optional[float] = None,
                 partition: Optional["Partition"] = None,
                 sample_batch_size: Optional[int] = None, autotune: bool = False) -> None:
        self.ctx = ctx
        iv = np.ascontiguousarray(np.asarray(intervals, dtype=np.float64).reshape(-1, 2))
        lmn = np.asfortranarray(np.asarray(lmn, dtype=np.float32))
        if lmn.ndim != 2 or lmn.shape[1] != 3:
            raise _capi._EXC[3]("lmn must be (nPixel, 3)")
        self.n_antenna, self.n_beam = int(n_antenna), int(n_beam)
        self.n_levels, self.n_pixel = iv.shape[0], lmn.shape[0]
        self.compute, self.imager = compute, imager
        h = C.c_void_p()
        check(lib().bipp_synthesis_create(
            ctx.handle, COMPUTE[compute], IMAGERS[imager], self.n_antenna, self.n_beam,
            self.n_levels, iv.ctypes.data_as(C.POINTER(C.c_double)), FILTERS[filter_name],
            self.n_pixel, lmn.ctypes.data_as(C.POINTER(C.c_float)), self.n_pixel,
            1 if batched else 0, C.byref(h)))
        self._h = h
        if tolerance is not None:
            check(lib().bipp_synthesis_set_tolerance(h, float(tolerance)))
        if partition is not None:
            m, dims = partition._c_args()
            check(lib().bipp_synthesis_set_partition(h, m, dims))
        if sample_batch_size:
            check(lib().bipp_synthesis_set_sample_batch_size(h, int(sample_batch_size)))
        if autotune:
            check(lib().bipp_synthesis_set_autotune(h, 1))

    def nufft_tuning(self):
        """(shape of the last NUFFT batch as a dict, list of (candidate dict, ms) timed by
        this handle's last tuning)."""
        keys = ("lines_radix", "pass_radix", "pass_tw_shift", "spread_chunk")
        shape = (C.c_int32 * 4)()
        n = C.c_size_t()
        check(lib().bipp_synthesis_nufft_tuning(self._h, shape, 0, None, C.byref(n)))
        cand = np.zeros(5 * max(n.value, 1))
        check(lib().bipp_synthesis_nufft_tuning(self._h, shape, n.value,
                                                cand.ctypes.data_as(C.POINTER(C.c_double)),
                                                C.byref(n)))
        rows = [(dict(zip(keys, (int(x) for x in cand[5 * k:5 * k + 4]))), float(cand[5 * k + 4]))
                for k in range(n.value)]
        return dict(zip(keys, list(shape))), rows

    def nufft_grid(self) -> Tuple[int, int, int]:
        nf = (C.c_int32 * 3)()
        check(lib().bipp_synthesis_nufft_grid(self._h, nf))
        return tuple(nf)

    def nufft_partition(self):
        """(boxes per axis, boxes holding points) of the last NUFFT collect."""
        dims = (C.c_int32 * 3)()
        boxes = C.c_int32()
        check(lib().bipp_synthesis_nufft_partition(self._h, dims, C.byref(boxes)))
        return tuple(dims), boxes.value

    def collect(self, wl, s, xyz, w=None) -> None:
        """Host arrays: wl (nS,) or scalar; s (nS, nBeam, nBeam) or (nBeam, nBeam), matrices
        in standard orientation (s[k][row, col]); xyz (nS, nAnt, 3) or (nAnt, 3);
        w (nS, nAnt, nBeam), (nAnt, nBeam) or None for identity beamforming."""
        s = np.asarray(s)
        if s.ndim == 2:
            s = s[None]
        n_s = s.shape[0]
        cdt = np.complex128 if s.dtype == np.complex128 else np.complex64
        prec = PRECISION["double" if cdt == np.complex128 else "single"]
        # column-major per sample == row-major of the transpose
        s_cm = np.ascontiguousarray(np.swapaxes(s, 1, 2).astype(cdt))
        xyz = np.asarray(xyz, dtype=np.float64)
        if xyz.ndim == 2:
            xyz = np.broadcast_to(xyz, (n_s,) + xyz.shape)
        x_cm = np.ascontiguousarray(np.swapaxes(xyz, 1, 2))
        wl = np.ascontiguousarray(np.broadcast_to(np.asarray(wl, dtype=np.float64), (n_s,)))
        w_cm, ldw, sw = None, 0, 0
        if w is not None:
            w = np.asarray(w)
            if w.ndim == 2:
                w = np.broadcast_to(w, (n_s,) + w.shape)
            w_cm = np.ascontiguousarray(np.swapaxes(w, 1, 2).astype(cdt))
            ldw, sw = self.n_antenna, self.n_antenna * self.n_beam
        dpp = C.POINTER(C.c_double)
        check(lib().bipp_synthesis_collect_host(
            self._h, n_s, wl.ctypes.data_as(dpp), prec, ptr(s_cm), self.n_beam,
            self.n_beam * self.n_beam, ptr(w_cm), ldw, sw, x_cm.ctypes.data_as(dpp),
            self.n_antenna, 3 * self.n_antenna))

    def collect_device(self, n_samples: int, wl_dev: int, s_dev: int, s_double: bool,
                       xyz_dev: int, w_dev: Optional[int] = None) -> None:
        """Device arrays in the batched C-ABI layout (see bipp_synthesis_collect)."""
        nb, na = self.n_beam, self.n_antenna
        check(lib().bipp_synthesis_collect(
            self._h, n_samples, wl_dev, PRECISION["double" if s_double else "single"], s_dev,
            nb, nb * nb, w_dev, na if w_dev else 0, na * nb if w_dev else 0, xyz_dev, na, 3 * na))

    def collect_graph(self, n_samples: int, wl_dev: int, s_dev: int, s_double: bool,
                      xyz_dev: int) -> None:
        """Device arrays (identity beamforming, steering imager): the chain is captured in a
        HIP graph on the first call for these pointers and replayed afterwards."""
        nb, na = self.n_beam, self.n_antenna
        check(lib().bipp_synthesis_collect_graph(
            self._h, n_samples, wl_dev, PRECISION["double" if s_double else "single"], s_dev,
            nb, nb * nb, xyz_dev, na, 3 * na))

    def graph_info(self) -> Tuple[int, int]:
        """(graph captures, graph launches) of collect_graph on this handle."""
        a, b = C.c_longlong(), C.c_longlong()
        check(lib().bipp_synthesis_graph_info(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def get(self, level: int, normalize: bool = True) -> np.ndarray:
        out = np.empty(self.n_pixel, dtype=np.float32)
        check(lib().bipp_synthesis_get(self._h, int(level), 1 if normalize else 0,
                                       out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def images(self, normalize: bool = True) -> np.ndarray:
        return np.stack([self.get(l, normalize) for l in range(self.n_levels)])

    def image_device(self) -> Tuple[int, str]:
        p = C.c_void_p()
        t = C.c_int()
        check(lib().bipp_synthesis_image_device(self._h, C.byref(p), C.byref(t)))
        return p.value, ("double" if t.value == 1 else "single")

    @property
    def num_collected(self) -> int:
        n = C.c_size_t()
        check(lib().bipp_synthesis_num_collected(self._h, C.byref(n)))
        return n.value

    def reset(self) -> None:
        check(lib().bipp_synthesis_reset(self._h))

    def timings(self, reset: bool = False) -> Dict[str, float]:
        """GPU milliseconds per stage accumulated over collect calls (HIP events)."""
        ms = np.zeros(9)
        check(lib().bipp_synthesis_timings(self._h, ms.ctypes.data_as(C.POINTER(C.c_double)),
                                           1 if reset else 0))
        keys = ("eigensolve_ms", "prep_ms", "image_ms", "batches", "nufft_vv_ms",
                "nufft_sort_ms", "nufft_spread_ms", "nufft_fft_ms", "nufft_interp_ms")
        return {k: float(v) for k, v in zip(keys, ms)}

    def last_eigenvalues(self):
        """(eigenvalues (nS, nBeam) float64, kept-beam counts (nS,)) of every sample of the
        last collect call."""
        n = C.c_size_t()
        check(lib().bipp_synthesis_last_eigenvalues(self._h, 0, None, None, C.byref(n)))
        d = np.zeros((n.value, self.n_beam), dtype=np.float64)
        nr = np.zeros(n.value, dtype=np.int32)
        check(lib().bipp_synthesis_last_eigenvalues(
            self._h, n.value, d.ctypes.data_as(C.POINTER(C.c_double)),
            nr.ctypes.data_as(C.POINTER(C.c_int32)), None))
        return d, nr

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and not sys.is_finalizing():
            lib().bipp_synthesis_destroy(h)
            self._h = None


from .synthesis import EigenpairImager, check_selection, image_synthesis  # noqa: E402,F401
