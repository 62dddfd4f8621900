"""In-memory datasets, image properties and image data.

The interfaces of include/bipp/dataset.hpp:14-41, image_prop.hpp:16-33 and image_data.hpp:16-37,
with in-memory implementations shaped on the file classes of the pybind module
(DatasetFile, ImagePropFile, ImageDataFile, python/pybipp.cpp:99-171,347-460).  The reference
stores these in HDF5; what a dataset holds is kept at the file's precision -- eigenvectors
complex64, eigenvalues, uvw, wavelength and scale float32 (src/dataset_file.cpp:28-73,
204-249) -- so synthesis sees the same rounded values it would read back from a file.
"""
from __future__ import annotations

import abc
from typing import Dict, List, Union

import numpy as np

from . import _capi

MetaType = Union[int, float, List[float]]


def _bad(msg: str) -> Exception:
    return _capi._EXC[3](msg)


class Dataset(abc.ABC):
    """bipp::Dataset (include/bipp/dataset.hpp:14-41)."""

    @abc.abstractmethod
    def description(self) -> str: ...

    @abc.abstractmethod
    def num_samples(self) -> int: ...

    @abc.abstractmethod
    def num_antenna(self) -> int: ...

    @abc.abstractmethod
    def num_beam(self) -> int: ...

    @abc.abstractmethod
    def eig_vec(self, index: int) -> np.ndarray:
        """(nAntenna, nBeam) complex64."""

    @abc.abstractmethod
    def eig_val(self, index: int) -> np.ndarray:
        """(nBeam,) float32."""

    @abc.abstractmethod
    def uvw(self, index: int) -> np.ndarray:
        """(nAntenna^2, 3) float32, metres, entry i + nAntenna j for antennas (i, j)."""

    @abc.abstractmethod
    def wl(self, index: int) -> float: ...

    @abc.abstractmethod
    def time_stamp(self, index: int) -> float: ...

    @abc.abstractmethod
    def scale(self, index: int) -> float: ...

    @abc.abstractmethod
    def ra_deg(self) -> float: ...

    @abc.abstractmethod
    def dec_deg(self) -> float: ...


class InMemoryDataset(Dataset):
    """A dataset held in host memory, written sample by sample as DatasetFile.write
    (include/bipp/dataset_file.hpp:48, python/pybipp.cpp:151-164)."""

    def __init__(self, description: str, n_antenna: int, n_beam: int, ra_deg: float = 0.0,
                 dec_deg: float = 0.0) -> None:
        self._desc = str(description)
        self._n_ant, self._n_beam = int(n_antenna), int(n_beam)
        self._ra, self._dec = float(ra_deg), float(dec_deg)
        self._v: List[np.ndarray] = []
        self._d: List[np.ndarray] = []
        self._uvw: List[np.ndarray] = []
        self._wl: List[np.float32] = []
        self._t: List[np.float32] = []
        self._sc: List[np.float32] = []

    @staticmethod
    def create(description: str, n_antenna: int, n_beam: int, ra_deg: float = 0.0,
               dec_deg: float = 0.0) -> "InMemoryDataset":
        return InMemoryDataset(description, n_antenna, n_beam, ra_deg, dec_deg)

    def write(self, time_stamp: float, wl: float, scale: float, v, d, uvw) -> None:
        v = np.asarray(v)
        d = np.asarray(d)
        uvw = np.asarray(uvw)
        if v.shape != (self._n_ant, self._n_beam):
            raise _bad(f"v must be ({self._n_ant}, {self._n_beam}), got {v.shape}")
        if d.shape != (self._n_beam,):
            raise _bad(f"d must be ({self._n_beam},), got {d.shape}")
        if uvw.shape != (self._n_ant * self._n_ant, 3):
            raise _bad(f"uvw must be ({self._n_ant ** 2}, 3), got {uvw.shape}")
        self._v.append(np.asfortranarray(v, dtype=np.complex64))
        self._d.append(np.ascontiguousarray(d, dtype=np.float32))
        self._uvw.append(np.asfortranarray(uvw, dtype=np.float32))
        self._wl.append(np.float32(wl))
        self._t.append(np.float32(time_stamp))
        self._sc.append(np.float32(scale))

    def _check(self, index: int) -> int:
        if not 0 <= index < len(self._v):
            raise _bad(f"sample index {index} out of range")
        return index

    def description(self) -> str:
        return self._desc

    def num_samples(self) -> int:
        return len(self._v)

    def num_antenna(self) -> int:
        return self._n_ant

    def num_beam(self) -> int:
        return self._n_beam

    def eig_vec(self, index: int) -> np.ndarray:
        return self._v[self._check(index)].copy(order="F")

    def eig_val(self, index: int) -> np.ndarray:
        return self._d[self._check(index)].copy()

    def uvw(self, index: int) -> np.ndarray:
        return self._uvw[self._check(index)].copy(order="F")

    def wl(self, index: int) -> float:
        return float(self._wl[self._check(index)])

    def time_stamp(self, index: int) -> float:
        return float(self._t[self._check(index)])

    def scale(self, index: int) -> float:
        return float(self._sc[self._check(index)])

    def ra_deg(self) -> float:
        return self._ra

    def dec_deg(self) -> float:
        return self._dec


class ImageProp(abc.ABC):
    """bipp::ImageProp (include/bipp/image_prop.hpp:16-33)."""

    @abc.abstractmethod
    def pixel_lmn(self) -> np.ndarray:
        """(width * height, 3) float32 direction cosines."""

    @abc.abstractmethod
    def meta_data(self) -> Dict[str, MetaType]: ...

    @abc.abstractmethod
    def set_meta(self, name: str, value: MetaType) -> None: ...

    @abc.abstractmethod
    def width(self) -> int: ...

    @abc.abstractmethod
    def height(self) -> int: ...

    @abc.abstractmethod
    def fov_deg(self) -> float: ...


class InMemoryImageProp(ImageProp):
    """Pixel grid and metadata in memory, created as ImagePropFile.create
    (include/bipp/image_prop_file.hpp:22, python/pybipp.cpp:374-387)."""

    def __init__(self, height: int, width: int, fov_deg: float, lmn) -> None:
        lmn = np.asarray(lmn)
        if lmn.shape != (int(height) * int(width), 3):
            raise _bad(f"lmn must be ({int(height) * int(width)}, 3), got {lmn.shape}")
        self._h, self._w, self._fov = int(height), int(width), float(fov_deg)
        self._lmn = np.asfortranarray(lmn, dtype=np.float32)
        self._meta: Dict[str, MetaType] = {}

    @staticmethod
    def create(height: int, width: int, fov_deg: float, lmn) -> "InMemoryImageProp":
        return InMemoryImageProp(height, width, fov_deg, lmn)

    def pixel_lmn(self) -> np.ndarray:
        return self._lmn.copy(order="F")

    def meta_data(self) -> Dict[str, MetaType]:
        return dict(self._meta)

    def set_meta(self, name: str, value: MetaType) -> None:
        self._meta[str(name)] = value

    def width(self) -> int:
        return self._w

    def height(self) -> int:
        return self._h

    def fov_deg(self) -> float:
        return self._fov


class ImageData(abc.ABC):
    """bipp::ImageData (include/bipp/image_data.hpp:16-37)."""

    @abc.abstractmethod
    def tags(self) -> List[str]: ...

    def num_tags(self) -> int:
        return len(self.tags())

    @abc.abstractmethod
    def get(self, tag: str) -> np.ndarray:
        """(height, width) float32, column-major as ImageDataFile.get returns it."""

    @abc.abstractmethod
    def set(self, tag: str, image) -> None: ...

    @abc.abstractmethod
    def width(self) -> int: ...

    @abc.abstractmethod
    def height(self) -> int: ...

    @abc.abstractmethod
    def fov_deg(self) -> float: ...

    @abc.abstractmethod
    def ra_deg(self) -> float: ...

    @abc.abstractmethod
    def dec_deg(self) -> float: ...


class InMemoryImageData(ImageData):
    """Images by tag, in memory (the ImageDataFile of python/pybipp.cpp:411-452)."""

    def __init__(self, height: int, width: int, fov_deg: float, ra_deg: float,
                 dec_deg: float) -> None:
        self._h, self._w = int(height), int(width)
        self._fov, self._ra, self._dec = float(fov_deg), float(ra_deg), float(dec_deg)
        self._img: Dict[str, np.ndarray] = {}

    @staticmethod
    def create(height: int, width: int, fov_deg: float, ra_deg: float,
               dec_deg: float) -> "InMemoryImageData":
        return InMemoryImageData(height, width, fov_deg, ra_deg, dec_deg)

    def tags(self) -> List[str]:
        return list(self._img)

    def get(self, tag: str) -> np.ndarray:
        if tag not in self._img:
            raise _bad(f"no image with tag {tag!r}")
        return self._img[tag].copy(order="F")

    def set(self, tag: str, image) -> None:
        img = np.asarray(image, dtype=np.float32)
        if img.size != self._h * self._w:
            raise _bad(f"image must hold {self._h * self._w} pixels, got {img.size}")
        # pixel q of the flat image is entry (q mod height, q / height) of the column-major array
        self._img[str(tag)] = np.asfortranarray(img.reshape(self._h, self._w, order="F"))

    def width(self) -> int:
        return self._w

    def height(self) -> int:
        return self._h

    def fov_deg(self) -> float:
        return self._fov

    def ra_deg(self) -> float:
        return self._ra

    def dec_deg(self) -> float:
        return self._dec
