"""The C-ABI library loads and exports what include/bipp_hip.h declares (no GPU needed)."""
import ctypes
import subprocess

import numpy as np
import pytest

import bipp_amd
from bipp_amd import _capi


def test_header_symbols_exported():
    syms = _capi.header_symbols()
    assert len(syms) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", _capi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_binding_covers_header():
    assert set(_capi.header_symbols()) == set(_capi._SIGS)
    lib = _capi.lib()
    for name in _capi._SIGS:
        assert getattr(lib, name) is not None


def test_version_and_config():
    assert bipp_amd.lib().bipp_version().decode().startswith("1.0.0")
    assert bipp_amd.config.rocm and not bipp_amd.config.cuda


def test_error_codes_match_reference_enum():
    # numeric values of include/bipp/errors.h are part of the interface
    assert _capi.ERROR_NAMES[6] == "EigensolverError"
    assert issubclass(bipp_amd.EigensolverError, bipp_amd.GenericError)
    assert bipp_amd.EigensolverError("x").error_code() == 6


def test_cpu_context_refused():
    with pytest.raises(bipp_amd.InvalidParameterError):
        bipp_amd.Context("CPU")


def test_null_handles_rejected():
    lib = _capi.lib()
    assert lib.bipp_ctx_synchronize(None) == 5
    assert "null" in lib.bipp_last_error().decode()
    assert lib.bipp_synthesis_reset(None) == 5


def test_options_mirror():
    opt = bipp_amd.NufftSynthesisOptions()
    assert opt.tolerance == pytest.approx(1e-3) and opt.precision == "single"
    opt.set_precision("double").set_tolerance(1e-5).set_normalize_image(False)
    assert opt.precision == "double" and opt.compute_mode() == "fp64"
    with pytest.raises(bipp_amd.InvalidParameterError):
        opt.set_precision("half")
    assert bipp_amd.Partition.grid([2, 3, 4]).dimensions == (2, 3, 4)
    assert bipp_amd.communicator.world().size == 1


def test_levels_match_oracle():
    from oracle import bipp_oracle as O
    d = np.abs(np.random.default_rng(0).lognormal(0, 2, size=300))
    a = bipp_amd.infer_intervals(4, 1.0, "log", 0.0, float("inf"), d)
    b = O.infer_intervals(4, 1.0, "log", 0.0, float("inf"), d)
    assert np.array_equal(a, b)
    f = bipp_amd.Filter("sqrt", 0.5, 10.0)
    x = np.array([0.25, 4.0, 16.0])
    assert np.array_equal(f(x), O.level_filter("sqrt", 0.5, 10.0, x))
