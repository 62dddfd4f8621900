"""ctypes binding of libbipp_amd.so (the C-ABI declared in include/bipp_hip.h).

The library is loaded on first use; a missing library raises ImportError naming the
build command -- there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Dict

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libbipp_amd.so")
HEADER = os.path.join(ROOT, "include", "bipp_hip.h")

# BippError (include/bipp/errors.h)
BIPP_SUCCESS = 0
ERROR_NAMES = {
    1: "GenericError", 2: "InternalError", 3: "InvalidParameterError",
    4: "InvalidPointerError", 5: "InvalidHandleError", 6: "EigensolverError",
    7: "NotImplementedError", 8: "GPUError", 9: "GPUSupportError", 10: "GPUBlasError",
    11: "InvalidAllocatorFunctionError", 12: "MPIError", 13: "HDF5Error", 14: "FileError",
}

# enums (include/bipp/enums.h)
PU = {"AUTO": 0, "CPU": 1, "GPU": 2}
PRECISION = {"single": 0, "double": 1}
COMPUTE = {"fp64": 0, "fp32": 1, "mixed": 2}
FILTERS = {"lsq": 0, "std": 1, "sqrt": 2, "inv": 3, "inv_sqrt": 4}
IMAGERS = {"steering": 0, "direct": 1, "nufft": 2}
PARTITION = {"auto": 0, "none": 1, "grid": 2}


class GenericError(RuntimeError):
    """Base of the library's exceptions (include/bipp/exceptions.hpp:16-140)."""

    code = 1

    def error_code(self) -> int:
        return self.code


_EXC: Dict[int, type] = {}
for _code, _name in ERROR_NAMES.items():
    if _code == 1:
        _EXC[_code] = GenericError
        continue
    _EXC[_code] = type(_name, (GenericError,), {"code": _code})
globals().update({cls.__name__: cls for cls in _EXC.values()})

_lib = None

sz = C.c_size_t
vp = C.c_void_p
i32 = C.c_int32
dp = C.POINTER(C.c_double)
fp = C.POINTER(C.c_float)

_SIGS = {
    "bipp_last_error": (C.c_char_p, []),
    "bipp_version": (C.c_char_p, []),
    "bipp_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "bipp_ctx_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "bipp_ctx_destroy": (C.c_int, [vp]),
    "bipp_ctx_processing_unit": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "bipp_ctx_stream": (C.c_int, [vp, C.POINTER(vp)]),
    "bipp_ctx_synchronize": (C.c_int, [vp]),
    "bipp_eigh_gram_s": (C.c_int, [C.c_float, sz, sz, vp, sz, vp, sz, fp, sz, fp, vp, sz,
                                   C.POINTER(sz), fp]),
    "bipp_eigh_gram_d": (C.c_int, [C.c_double, sz, sz, vp, sz, vp, sz, dp, sz, dp, vp, sz,
                                   C.POINTER(sz), dp]),
    "bipp_eigh_s": (C.c_int, [C.c_float, sz, sz, vp, sz, vp, sz, fp, vp, sz, C.POINTER(sz), fp]),
    "bipp_eigh_d": (C.c_int, [C.c_double, sz, sz, vp, sz, vp, sz, dp, vp, sz, C.POINTER(sz), dp]),
    "bipp_eigh_gram_batched": (C.c_int, [vp, sz, sz, sz, C.c_int, C.c_int, C.c_int, vp, vp, sz,
                                         sz, vp, sz, sz, vp, sz, sz, vp, sz, vp, sz, sz, vp, vp]),
    "bipp_eigh_gram_batched_noalloc": (C.c_int, [vp, sz, sz, sz, C.c_int, C.c_int, C.c_int, vp,
                                                 vp, sz, sz, vp, sz, sz, vp, sz, sz, vp, sz, vp,
                                                 sz, sz, vp, vp]),
    "bipp_synthesis_create": (C.c_int, [vp, C.c_int, C.c_int, sz, sz, sz, dp, C.c_int, sz, fp,
                                        sz, C.c_int, C.POINTER(vp)]),
    "bipp_synthesis_destroy": (C.c_int, [vp]),
    "bipp_synthesis_collect": (C.c_int, [vp, sz, vp, C.c_int, vp, sz, sz, vp, sz, sz, vp, sz,
                                         sz]),
    "bipp_synthesis_collect_graph": (C.c_int, [vp, sz, vp, C.c_int, vp, sz, sz, vp, sz, sz]),
    "bipp_synthesis_graph_info": (C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "bipp_synthesis_collect_host": (C.c_int, [vp, sz, dp, C.c_int, vp, sz, sz, vp, sz, sz, dp,
                                              sz, sz]),
    "bipp_synthesis_get": (C.c_int, [vp, sz, C.c_int, fp]),
    "bipp_synthesis_image_device": (C.c_int, [vp, C.POINTER(vp), C.POINTER(C.c_int)]),
    "bipp_synthesis_num_collected": (C.c_int, [vp, C.POINTER(sz)]),
    "bipp_synthesis_reset": (C.c_int, [vp]),
    "bipp_synthesis_last_eigenvalues": (C.c_int, [vp, sz, dp, C.POINTER(i32), C.POINTER(sz)]),
    "bipp_synthesis_timings": (C.c_int, [vp, dp, C.c_int]),
    "bipp_synthesis_set_tolerance": (C.c_int, [vp, C.c_double]),
    "bipp_synthesis_nufft_grid": (C.c_int, [vp, C.POINTER(i32)]),
    "bipp_synthesis_set_partition": (C.c_int, [vp, C.c_int, C.POINTER(sz)]),
    "bipp_synthesis_nufft_partition": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "bipp_synthesis_set_sample_batch_size": (C.c_int, [vp, sz]),
    "bipp_synthesis_set_autotune": (C.c_int, [vp, C.c_int]),
    "bipp_synthesis_nufft_tuning": (C.c_int, [vp, C.POINTER(i32), sz, dp, C.POINTER(sz)]),
    "bipp_nufft_tuning_reset": (C.c_int, []),
    "bipp_imager_create": (C.c_int, [vp, C.c_int, C.c_int, sz, sz, sz, sz, fp, sz, C.POINTER(vp)]),
    "bipp_imager_destroy": (C.c_int, [vp]),
    "bipp_imager_set_tolerance": (C.c_int, [vp, C.c_double]),
    "bipp_imager_set_partition": (C.c_int, [vp, C.c_int, C.POINTER(sz)]),
    "bipp_imager_set_sample_batch_size": (C.c_int, [vp, sz]),
    "bipp_imager_add": (C.c_int, [vp, C.c_float, C.c_float, fp, sz, vp, sz, fp, sz]),
    "bipp_imager_get": (C.c_int, [vp, sz, fp]),
    "bipp_imager_num_added": (C.c_int, [vp, C.POINTER(sz)]),
    "bipp_imager_nufft_partition": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32)]),
    "bipp_virtual_vis": (C.c_int, [vp, C.c_int, sz, sz, sz, vp, sz, sz, vp, sz, vp, sz]),
    "bipp_nuft_sum": (C.c_int, [vp, C.c_int, C.c_double, sz, vp, vp, vp, vp, sz, vp, vp, vp,
                                vp]),
    "bipp_nuft_sum_mixed": (C.c_int, [vp, C.c_double, sz, vp, vp, vp, vp, sz, vp, vp, vp, vp]),
    "bipp_gram": (C.c_int, [vp, C.c_int, sz, sz, vp, vp, sz, sz, C.c_double, vp, sz, sz]),
}


def header_symbols():
    """Function names declared in include/bipp_hip.h."""
    with open(HEADER) as f:
        txt = f.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[A-Za-z_]+\s*\*?\s*(bipp_[a-z0-9_]+)\s*\(",
                                 txt, re.M)))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run `python -m bipp_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(code: int) -> None:
    if code != BIPP_SUCCESS:
        msg = lib().bipp_last_error().decode(errors="replace")
        raise _EXC.get(code, GenericError)(msg or ERROR_NAMES.get(code, f"error {code}"))


def call(name: str, *args) -> None:
    check(getattr(lib(), name)(*args))


def ptr(a) -> int:
    """Address of a numpy array or a torch tensor (device or host)."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data
