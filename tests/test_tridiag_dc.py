"""The divide-and-conquer tridiagonal eigensolver of the eigensolver kernel, run on the CPU.

tests/csrc/dc_harness.cpp compiles bipp_amd/csrc/tridiag_dc.hpp for the host with one thread
per workgroup thread and a barrier for the workgroup barrier (the GPU's decomposition and
synchronisation), under AddressSanitizer and UndefinedBehaviorSanitizer; its scratch arrays
start as NaN so that a read of memory the solver never wrote shows in the results.  The
eigenpairs are checked against LAPACK's tridiagonal solver (scipy.linalg.eigh_tridiagonal):
eigenvalues, residual |T Z - Z diag(d)| and orthogonality |Z^T Z - I|, over every size bucket
of the kernel (64, 128, 256, 512) and matrices with clustered, repeated and split spectra.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.linalg import eigh_tridiagonal, hessenberg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("dc") / "dc_harness")
    subprocess.run([gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-pthread",
                    os.path.join(ROOT, "tests", "csrc", "dc_harness.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def run(exe, tmp_path, d, e, nmax, single=False):
    n = len(d)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([nmax, n, int(single)], np.int32).tofile(f)
        np.asarray(d, np.float64).tofile(f)
        np.asarray(e, np.float64).tofile(f)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    o = np.fromfile(fout, np.float64)
    return o[:n], o[n:].reshape(n, n, order="F")


def check(exe, tmp_path, d, e, nmax, single=False):
    lam, Z = run(exe, tmp_path, d, e, nmax, single)
    n = len(d)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    ref = eigh_tridiagonal(d, e, eigvals_only=True) if n > 1 else np.asarray(d, float)
    scale = max(float(np.max(np.abs(ref))), 1e-300)
    tol = 2e-6 if single else 1e-13
    assert np.all(np.diff(lam) >= 0), "eigenvalues ascending"
    assert np.max(np.abs(lam - ref)) <= tol * scale
    assert np.linalg.norm(T @ Z - Z * lam[None, :]) <= tol * scale * np.sqrt(n)
    assert np.linalg.norm(Z.T @ Z - np.eye(n)) <= tol * np.sqrt(n)


@pytest.mark.parametrize("n,nmax", [(1, 64), (2, 64), (3, 64), (37, 64), (64, 64), (100, 128),
                                    (256, 256), (300, 512)])
def test_random(harness, tmp_path, n, nmax):
    rng = np.random.default_rng(n)
    check(harness, tmp_path, rng.standard_normal(n), rng.standard_normal(n - 1), nmax)


def _cases(n, rng):
    w = np.abs(np.arange(-(n // 2), n - n // 2)).astype(float)
    glued_e = rng.standard_normal(n - 1)
    glued_e[::16] = 1e-18
    A = rng.standard_normal((n, 4)) @ np.diag([100.0, 30.0, 10.0, 3.0])
    H = hessenberg(A @ A.T + 1e-6 * np.eye(n))
    return {
        "repeated diagonal, tiny coupling": (np.ones(n), np.full(n - 1, 1e-9)),
        "diagonal (no coupling)": (np.arange(n, dtype=float)[::-1].copy(), np.zeros(n - 1)),
        "Wilkinson": (w, np.ones(n - 1)),
        "zero diagonal Toeplitz": (np.zeros(n), np.ones(n - 1)),
        "glued blocks": (rng.standard_normal(n), glued_e),
        "rank 4 + 1e-6 I (visibility-like)": (np.diag(H).copy(), np.diag(H, -1).copy()),
        "all zero": (np.zeros(n), np.zeros(n - 1)),
    }


@pytest.mark.parametrize("name", list(_cases(8, np.random.default_rng(0))))
def test_hard_spectra(harness, tmp_path, name):
    d, e = _cases(200, np.random.default_rng(7))[name]
    check(harness, tmp_path, d, e, 256)


def test_single_precision(harness, tmp_path):
    rng = np.random.default_rng(3)
    check(harness, tmp_path, rng.standard_normal(128), rng.standard_normal(127), 128, single=True)
