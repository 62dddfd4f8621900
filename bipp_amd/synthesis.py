"""image_synthesis on in-memory datasets: the reference's stage B on the GPU.

bipp::image_synthesis (include/bipp/image_synthesis.hpp:160-163, src/image_synthesis.cpp:28-182,
pybind form python/pybipp.cpp:174-196): checks and sorts the selection, gathers each image's
scaled eigenvalues per sample, feeds every selected sample's eigenpairs and uvw to the
eigenpair imager of libbipp_amd (virtual visibilities and the type-3 NUFFT or direct sum on
the GPU, src/nufft_synthesis.cpp:74-170) and normalises by the number of samples.  The
images go to an ImageData object instead of an HDF5 file.
"""
from __future__ import annotations

import ctypes as C
import sys
from typing import Dict, Iterable, List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _capi
from ._capi import COMPUTE, IMAGERS, check, lib
from .dataset import Dataset, ImageData, ImageProp, InMemoryImageData


def _bad(msg: str) -> Exception:
    return _capi._EXC[3](msg)


class EigenpairImager:
    """Stage-B imager of libbipp_amd (bipp_imager_* in include/bipp_hip.h): add(sample) /
    get(image), the NUFFTInterface::add / get_image pair of src/nufft_interface.hpp:13-21 with
    eigenpairs instead of host-formed virtual visibilities."""

    def __init__(self, ctx, n_antenna: int, n_beam: int, n_images: int, lmn,
                 compute: str = "mixed", imager: str = "nufft", tolerance: float = 1e-5,
                 partition=None, sample_batch_size: Optional[int] = None) -> None:
        if imager not in ("nufft", "direct"):
            raise _bad(f"the eigenpair imager runs 'nufft' or 'direct', not {imager!r}")
        if compute not in COMPUTE:
            raise _bad(f"unknown compute mode {compute!r}")
        lmn = np.asfortranarray(np.asarray(lmn, dtype=np.float32))
        if lmn.ndim != 2 or lmn.shape[1] != 3:
            raise _bad("lmn must be (nPixel, 3)")
        self.ctx = ctx
        self.n_antenna, self.n_beam, self.n_images = int(n_antenna), int(n_beam), int(n_images)
        self.n_pixel = lmn.shape[0]
        h = C.c_void_p()
        check(lib().bipp_imager_create(ctx.handle, COMPUTE[compute], IMAGERS[imager],
                                       self.n_antenna, self.n_beam, self.n_images, self.n_pixel,
                                       lmn.ctypes.data_as(C.POINTER(C.c_float)), self.n_pixel,
                                       C.byref(h)))
        self._h = h
        check(lib().bipp_imager_set_tolerance(h, float(tolerance)))
        if partition is not None:
            m, dims = partition._c_args()
            check(lib().bipp_imager_set_partition(h, m, dims))
        if sample_batch_size:
            check(lib().bipp_imager_set_sample_batch_size(h, int(sample_batch_size)))

    def add(self, wl: float, scale: float, uvw, v, d_scaled) -> None:
        """uvw (nAntenna^2, 3) metres; v (nAntenna, nBeam) complex64; d_scaled (nImages, nBeam)."""
        uvw = np.asfortranarray(np.asarray(uvw, dtype=np.float32))
        v = np.asfortranarray(np.asarray(v, dtype=np.complex64))
        d = np.ascontiguousarray(np.asarray(d_scaled, dtype=np.float32))
        na, nb = self.n_antenna, self.n_beam
        if uvw.shape != (na * na, 3) or v.shape != (na, nb) or d.shape != (self.n_images, nb):
            raise _bad("sample arrays do not match the imager's antenna, beam and image counts")
        fpp = C.POINTER(C.c_float)
        check(lib().bipp_imager_add(self._h, float(wl), float(scale), uvw.ctypes.data_as(fpp),
                                    na * na, v.ctypes.data, na, d.ctypes.data_as(fpp), nb))

    def get(self, index: int) -> np.ndarray:
        out = np.empty(self.n_pixel, dtype=np.float32)
        check(lib().bipp_imager_get(self._h, int(index), out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    @property
    def num_added(self) -> int:
        n = C.c_size_t()
        check(lib().bipp_imager_num_added(self._h, C.byref(n)))
        return n.value

    def nufft_partition(self):
        dims = (C.c_int32 * 3)()
        boxes = C.c_int32()
        check(lib().bipp_imager_nufft_partition(self._h, dims, C.byref(boxes)))
        return tuple(dims), boxes.value

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h is not None and h.value and not sys.is_finalizing():
            lib().bipp_imager_destroy(h)
            self._h = None


SelectionType = Dict[str, Union[Dict[int, Sequence[float]], Iterable[Tuple[int, Sequence[float]]]]]


def check_selection(selection: SelectionType, n_samples: int, n_beam: int):
    """The selection checks of src/image_synthesis.cpp:37-104 and python/pybipp.cpp:181-190.
    Returns (image names, sorted sample ids, dScaled float32 (nImages, nSamples, nBeam))."""
    if not selection:
        raise _bad("Empty selection")
    names: List[str] = []
    per_image: List[List[Tuple[int, np.ndarray]]] = []
    for name, samples in selection.items():
        name = str(name)
        for ch, what in (("/", "'/'"), (".", "'.'"), (" ", "spaces")):
            if ch in name:
                raise _bad(f"Selection name must not contain {what}")
        items = list(samples.items()) if isinstance(samples, dict) else list(samples)
        if not items:
            raise _bad("Empty selection")
        conv = []
        for sid, vals in items:
            arr = np.asarray(vals, dtype=np.float32).ravel()
            if arr.size != n_beam:
                raise _bad("Number of eigenvalues in selection does not match number of beams.")
            conv.append((int(sid), arr))
        conv.sort(key=lambda x: x[0])
        ids = [i for i, _ in conv]
        if any(a == b for a, b in zip(ids, ids[1:])):
            raise _bad("Selection sample id duplicate")
        if ids[0] < 0 or ids[-1] >= n_samples:
            raise _bad("Selection sample id exceeds number of dataset samples")
        names.append(name)
        per_image.append(conv)
    ids0 = [i for i, _ in per_image[0]]
    for conv in per_image[1:]:
        if [i for i, _ in conv] != ids0:
            raise _bad("Selection must contain the same sample ids for all images.")
    d_scaled = np.stack([np.stack([a for _, a in conv]) for conv in per_image])
    return names, ids0, d_scaled


def image_synthesis(ctx, opt, dataset: Dataset, selection: SelectionType, image_prop: ImageProp,
                    image_data: Optional[ImageData] = None) -> ImageData:
    """bipp.image_synthesis(ctx, opt, dataset, selection, image_prop) with the images returned
    in an ImageData (created in memory unless given) instead of written to a file.

    opt: NufftSynthesisOptions -- tolerance, precision ("single": float NUFFT grid with double
    virtual visibilities and point coordinates; "double": double throughout), sample batch
    size, uvw partition and image normalisation act as in the reference; `apply_scaling` is
    not read there either (src/nufft_synthesis.cpp:138,160)."""
    names, ids, d_scaled = check_selection(selection, dataset.num_samples(), dataset.num_beam())
    n_ant, n_beam = dataset.num_antenna(), dataset.num_beam()
    lmn = np.asarray(image_prop.pixel_lmn(), dtype=np.float32)
    if lmn.shape != (image_prop.width() * image_prop.height(), 3):
        raise _bad("pixel_lmn does not match the image size")
    imager = opt.imager if opt.imager in ("nufft", "direct") else "nufft"
    im = EigenpairImager(ctx, n_ant, n_beam, len(names), lmn, compute=opt.compute_mode(),
                         imager=imager, tolerance=opt.tolerance,
                         partition=opt.local_uvw_partition,
                         sample_batch_size=opt.sample_batch_size)
    for k, sid in enumerate(ids):
        im.add(dataset.wl(sid), dataset.scale(sid), dataset.uvw(sid), dataset.eig_vec(sid),
               d_scaled[:, k, :])
    out = image_data or InMemoryImageData(image_prop.height(), image_prop.width(),
                                          image_prop.fov_deg(), dataset.ra_deg(),
                                          dataset.dec_deg())
    norm = np.float32(1.0) / np.float32(len(ids))
    for i, name in enumerate(names):
        img = im.get(i)
        if opt.normalize_image:
            img = img * norm
        out.set(name, img)
    return out
