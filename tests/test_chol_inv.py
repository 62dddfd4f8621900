"""The blocked Cholesky factor and inverse of the eigensolver kernel, run on the CPU.

tests/csrc/ci_harness.cpp compiles bipp_amd/csrc/chol_inv.hpp for the host with one thread
per workgroup thread and a barrier for the workgroup barrier, under AddressSanitizer and
UndefinedBehaviorSanitizer; the strict upper triangle and the scratch array start as NaN.
W = L^-1 is checked against numpy's Cholesky factor inverted, W G W^H against I, the upper
triangle for not being written, and a matrix whose leading minor of order j + 1 is not
positive definite for the column LAPACK's zpotrf reports (info = j + 1).
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.linalg import lapack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("ci") / "ci_harness")
    subprocess.run([gxx, "-std=c++20", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-pthread",
                    os.path.join(ROOT, "tests", "csrc", "ci_harness.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def run(exe, tmp_path, G, nt, single=False):
    n = G.shape[0]
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    with open(fin, "wb") as f:
        np.array([nt, n, int(single)], np.int32).tofile(f)
        np.asfortranarray(G, np.complex128).ravel(order="F").tofile(f)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    raw = open(fout, "rb").read()
    return int(np.frombuffer(raw[:4], np.int32)[0]), \
        np.frombuffer(raw[4:], np.complex128).reshape(n, n, order="F")


def hpd(n, rng, cond=1e3):
    A = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Q, _ = np.linalg.qr(A)
    return (Q * np.logspace(0, np.log10(cond), n)) @ Q.conj().T


@pytest.mark.parametrize("n,nt", [(1, 256), (5, 256), (32, 512), (33, 256), (100, 256),
                                  (129, 512), (300, 512)])
def test_inverse_factor(harness, tmp_path, n, nt):
    rng = np.random.default_rng(n)
    G = hpd(n, rng)
    info, W = run(harness, tmp_path, G, nt)
    assert info == -1
    Wl = np.tril(W)
    Wref = np.linalg.inv(np.linalg.cholesky(G))
    assert np.linalg.norm(Wl - Wref) <= 1e-12 * np.linalg.norm(Wref)
    assert np.linalg.norm(Wl @ G @ Wl.conj().T - np.eye(n)) <= 1e-12 * np.sqrt(n)
    assert np.all(np.isnan(W[np.triu_indices(n, 1)])), "strict upper triangle untouched"


@pytest.mark.parametrize("n,bad", [(40, 35), (70, 3), (256, 200)])
def test_not_positive_definite(harness, tmp_path, n, bad):
    rng = np.random.default_rng(bad)
    G = hpd(n, rng)
    G[bad, bad] = -abs(G[bad, bad])
    _, info_ref = lapack.zpotrf(G, lower=1)
    info, _ = run(harness, tmp_path, G, 512 if n >= 256 else 256)
    assert info + 1 == info_ref


def test_single_precision(harness, tmp_path):
    rng = np.random.default_rng(9)
    G = hpd(150, rng, cond=1e2)
    info, W = run(harness, tmp_path, G, 512, single=True)
    Wref = np.linalg.inv(np.linalg.cholesky(G))
    assert info == -1
    assert np.linalg.norm(np.tril(W) - Wref) <= 1e-5 * np.linalg.norm(Wref)
