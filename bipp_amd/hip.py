"""Minimal HIP runtime access (device buffers, copies, events) through ctypes.

Uses the same libamdhip64 instance as libbipp_amd.so, so device pointers and streams
interoperate with the C-ABI without a second GPU runtime in the process.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import numpy as np

from . import _capi

_hip = None
H2D, D2H, D2D = 1, 2, 3


def hip():
    global _hip
    if _hip is None:
        _capi.lib()  # loads libamdhip64.so.7 as a dependency first
        L = C.CDLL("libamdhip64.so.7")
        L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        L.hipFree.argtypes = [C.c_void_p]
        L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        L.hipDeviceSynchronize.argtypes = []
        L.hipStreamSynchronize.argtypes = [C.c_void_p]
        L.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        L.hipEventDestroy.argtypes = [C.c_void_p]
        L.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        L.hipEventSynchronize.argtypes = [C.c_void_p]
        L.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        L.hipGetErrorString.restype = C.c_char_p
        L.hipGetErrorString.argtypes = [C.c_int]
        L.hipMemGetInfo.argtypes = [C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
        _hip = L
    return _hip


def _ok(err: int, what: str) -> None:
    if err != 0:
        raise _capi._EXC[8](f"{what}: {hip().hipGetErrorString(err).decode()}")


class DeviceArray:
    """A device allocation holding an array of a numpy dtype and shape."""

    def __init__(self, shape, dtype) -> None:
        self.shape = tuple(shape) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(self.nbytes, 1)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        _ok(hip().hipMemcpy(d.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return d

    @classmethod
    def zeros(cls, shape, dtype) -> "DeviceArray":
        d = cls(shape, dtype)
        _ok(hip().hipMemset(d.ptr, 0, max(d.nbytes, 1)), "hipMemset")
        return d

    def upload(self, a: np.ndarray) -> None:
        """Copy host data of the same dtype and byte size into this allocation."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("upload size differs from the allocation")
        synchronize()
        _ok(hip().hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")

    def numpy(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        synchronize()
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, D2H), "hipMemcpy D2H")
        return out

    def data_ptr(self) -> int:
        return self.ptr

    def free(self) -> None:
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self) -> None:
        import sys
        if getattr(self, "ptr", None) and not sys.is_finalizing():
            self.free()


def copy_device(dst: DeviceArray, src: DeviceArray, nbytes: int) -> None:
    """Device-to-device copy of `nbytes` (the runtime's copy kernel)."""
    _ok(hip().hipMemcpy(dst.ptr, src.ptr, nbytes, D2D), "hipMemcpy D2D")


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


def mem_info():
    free, total = C.c_size_t(), C.c_size_t()
    _ok(hip().hipMemGetInfo(C.byref(free), C.byref(total)), "hipMemGetInfo")
    return free.value, total.value


class Event:
    def __init__(self) -> None:
        e = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.e = e

    def record(self, stream: Optional[int] = None) -> None:
        _ok(hip().hipEventRecord(self.e, stream or None), "hipEventRecord")

    def synchronize(self) -> None:
        _ok(hip().hipEventSynchronize(self.e), "hipEventSynchronize")

    def elapsed_ms(self, end: "Event") -> float:
        t = C.c_float()
        _ok(hip().hipEventElapsedTime(C.byref(t), self.e, end.e), "hipEventElapsedTime")
        return float(t.value)
