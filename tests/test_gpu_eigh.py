"""Stage A parity on the GPU: flag/reduce, Gram, generalized eigensolve, un-beam.

Oracle: LAPACK ?hegv via scipy (src/host/lapack_api.hpp:112-158) composed as in
src/eigensolver.cpp:129-225.  Eigenvectors are compared through phase-invariant
quantities (V diag(d) V^H and the pencil residual), since eigenvectors of (nearly)
degenerate eigenvalues are not unique.
"""
import numpy as np
import pytest

import bipp_amd
from bipp_amd import synthetic
from bipp_amd.hip import DeviceArray
from oracle import bipp_oracle as O

pytestmark = pytest.mark.gpu


def vdv(v, d):
    return (v * d) @ v.conj().T


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def cfg1():
    return synthetic.make_workload(1, n_timesteps=3, grid=16)


@pytest.mark.parametrize("single", [False, True])
def test_eigh_gram_cfg1(ctx, cfg1, single):
    for smp in cfg1.samples:
        rd, cd = (np.float32, np.complex64) if single else (np.float64, np.complex128)
        n = smp.xyz.shape[0]
        print(f"eigh_gram cfg1 N={n} single={single}", flush=True)
        v, d, sc = bipp_amd.eigh_gram(rd(smp.wl), smp.s.astype(cd), np.eye(n, dtype=cd),
                                      smp.xyz.astype(rd))
        vo, do, nro, sco = O.eigh_gram(smp.wl, smp.s, None, smp.xyz, single=False)
        assert d.dtype == rd and v.dtype == cd
        tol = 2e-5 if single else 1e-11
        dmax = np.abs(do).max()
        assert np.abs(d - do).max() <= tol * dmax
        assert np.all(np.diff(d[:nro]) >= -tol * dmax)
        assert sc == pytest.approx(1.0 / (n * n), rel=2e-7)
        assert rel(vdv(v.astype(np.complex128), d), vdv(vo, do)) < (1e-4 if single else 1e-10)
        # pencil invariants: Z^H G Z = I, S Z = G Z diag(d)
        g = O.sinc_matrix(smp.xyz, smp.wl)
        z = v.astype(np.complex128)
        assert np.abs(z.conj().T @ g @ z - np.eye(n)).max() < (1e-4 if single else 1e-10)
        res = smp.s @ z - g @ z * d.astype(np.float64)
        assert np.abs(res).max() < (1e-4 if single else 1e-10) * dmax


@pytest.mark.parametrize("single", [False, True])
def test_eigh_gram_lofar_general_w(ctx, golden, single):
    """The reference's LOFAR fixture: beamforming W (55x55) with one all-zero column."""
    z = golden("lofar_input.npz")
    e = golden("lofar_expected.npz")
    tag = "f32" if single else "f64"
    rd, cd = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    for k in range(z["s"].shape[0]):
        print(f"eigh_gram lofar sample {k} single={single}", flush=True)
        v, d, sc = bipp_amd.eigh_gram(rd(z["wl"]), z["s"][k].astype(cd), z["w"][k].astype(cd),
                                      z["xyz"][k].astype(rd))
        nr = int(e[f"nr_{tag}"][k])
        assert nr == 54
        assert np.all(d[nr:] == 0) and np.all(v[:, nr:] == 0)
        assert sc == rd(e[f"scale_{tag}"][k])
        do, vo = e["d_f64"][k], e["v_f64"][k]
        dmax = np.abs(do).max()
        assert np.abs(d - do).max() <= (1e-4 if single else 1e-10) * dmax
        assert rel(vdv(v.astype(np.complex128), d.astype(np.float64)), vdv(vo, do)) < (
            1e-4 if single else 1e-9)


def test_flagged_antennas_reduce(ctx):
    wl_ = synthetic.make_workload(3, n_timesteps=2, n_channels=1, n_antenna=64)
    for smp in wl_.samples:
        n = smp.xyz.shape[0]
        v, d, sc = bipp_amd.eigh_gram(smp.wl, smp.s, np.eye(n), smp.xyz)
        vo, do, nro, sco = O.eigh_gram(smp.wl, smp.s, None, smp.xyz)
        assert nro == smp.active.sum() and np.all(d[nro:] == 0)
        assert np.all(v[~smp.active] == 0)
        assert np.abs(d - do).max() < 1e-10 * np.abs(do).max()
        assert sc == pytest.approx(sco, rel=1e-15)
        assert rel(vdv(v, d), vdv(vo, do)) < 1e-10


def test_all_zero_visibilities(ctx):
    n = 8
    v, d, sc = bipp_amd.eigh_gram(2.0, np.zeros((n, n), np.complex128), np.eye(n),
                                  np.random.default_rng(0).normal(size=(n, 3)))
    assert sc == 0 and np.all(d == 0) and np.all(v == 0)


def test_eigh_standard(ctx):
    r = np.random.default_rng(7)
    n = 40
    a = r.normal(size=(n, n)) + 1j * r.normal(size=(n, n))
    s = a @ a.conj().T
    s[5, :] = 0
    s[:, 5] = 0
    w = r.normal(size=(n, n)) + 1j * r.normal(size=(n, n))
    v, d, sc = bipp_amd.eigh(1.0, s, w)
    vo, do, nro, sco = O.eigh(1.0, s, w)
    assert nro == n - 1
    assert np.abs(d - do).max() < 1e-11 * np.abs(do).max()
    assert rel(vdv(v, d), vdv(vo, do)) < 1e-10
    assert sc == pytest.approx(sco)


def test_not_positive_definite_gram_raises(ctx):
    """A live beam with an all-zero weight column makes G exactly singular; LAPACK ?hegv
    reports info = n + 3 there, the reference throws EigensolverError
    (src/host/lapack_api.hpp:133)."""
    xyz = np.random.default_rng(1).normal(size=(6, 3)) * 10
    w = np.eye(6)
    w[:, 2] = 0
    s = np.eye(6, dtype=np.complex128) + 0.1
    with pytest.raises(np.linalg.LinAlgError):
        O.eigh_gram(2.0, s, w, xyz)
    with pytest.raises(bipp_amd.EigensolverError):
        bipp_amd.eigh_gram(2.0, s, w, xyz)


def test_batched_ragged_device(ctx):
    """Ragged batch (10..64 active of 64 antennas) in one persistent launch through the
    device-array entry point, against the oracle per sample."""
    r = np.random.default_rng(11)
    n, ns = 64, 24
    p0 = synthetic.disk_positions(r, n, 300.0)
    src, flux = synthetic.zenith_cap_sources(r, 12, 5.0)
    samples = []
    for k in range(ns):
        act = np.zeros(n, bool)
        act[r.choice(n, size=int(r.integers(10, n + 1)), replace=False)] = True
        wl = 2.0 + 0.1 * k
        s = synthetic.visibilities(p0, wl, src, flux, 0.1)
        s[~act] = 0
        s[:, ~act] = 0
        samples.append(synthetic.Sample(0, k, wl, p0, s, act))
    s_cm, x_cm, wls = synthetic.stack_samples(samples, np.complex64)
    s_d = DeviceArray.from_numpy(s_cm)
    x_d = DeviceArray.from_numpy(x_cm)
    wl_d = DeviceArray.from_numpy(wls)
    d_d = DeviceArray.zeros((ns, n), np.float64)
    v_d = DeviceArray.zeros((ns, n, n), np.complex128)
    nr_d = DeviceArray.zeros(ns, np.int32)
    sc_d = DeviceArray.zeros(ns, np.float64)
    print("batched ragged eigensolve", flush=True)
    bipp_amd.call("bipp_eigh_gram_batched", ctx.handle, ns, n, n, 0, 1, 1, wl_d.ptr, s_d.ptr, n,
                  n * n, None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, v_d.ptr, n, n * n, nr_d.ptr,
                  sc_d.ptr)
    ctx.synchronize()
    d = d_d.numpy()
    v = v_d.numpy().transpose(0, 2, 1)  # column-major per sample
    nr = nr_d.numpy()
    sc = sc_d.numpy()
    for k, smp in enumerate(samples):
        vo, do, nro, sco = O.eigh_gram(smp.wl, smp.s.astype(np.complex64), None, smp.xyz)
        assert nr[k] == nro
        assert sc[k] == pytest.approx(sco, rel=1e-15)
        assert np.abs(d[k] - do).max() < 1e-10 * np.abs(do).max()
        assert rel(vdv(v[k], d[k]), vdv(vo, do)) < 1e-10


@pytest.mark.parametrize("ragged,single", [(False, False), (True, False), (True, True)])
def test_batched_256_bucket(ctx, ragged, single):
    """256-antenna pencils (the configuration-4 size: in float64 the matrix-instruction
    phases — Cholesky and whitening tiles, divide-and-conquer products on 64 x 64 and 32 x 32
    tiles, the compact-WY back-transform; in float32 the vector-ALU forms), full or with
    180..256 active antennas, in one persistent launch, against the oracle per sample
    (LAPACK zhegv)."""
    r = np.random.default_rng(17 + int(ragged) + 2 * int(single))
    n, ns = 256, 6
    p0 = synthetic.disk_positions(r, n, 300.0)
    src, flux = synthetic.zenith_cap_sources(r, 24, 5.0)
    samples = []
    for k in range(ns):
        act = np.ones(n, bool)
        if ragged:
            act[:] = False
            act[r.choice(n, size=int(r.integers(180, n + 1)), replace=False)] = True
        wl = 2.0 + 0.05 * k
        s = synthetic.visibilities(p0, wl, src, flux, 0.1)
        s[~act] = 0
        s[:, ~act] = 0
        samples.append(synthetic.Sample(0, k, wl, p0, s, act))
    s_cm, x_cm, wls = synthetic.stack_samples(samples, np.complex64)
    s_d = DeviceArray.from_numpy(s_cm)
    x_d = DeviceArray.from_numpy(x_cm)
    wl_d = DeviceArray.from_numpy(wls)
    rd, cd = (np.float32, np.complex64) if single else (np.float64, np.complex128)
    prec = 0 if single else 1
    d_d = DeviceArray.zeros((ns, n), rd)
    v_d = DeviceArray.zeros((ns, n, n), cd)
    nr_d = DeviceArray.zeros(ns, np.int32)
    sc_d = DeviceArray.zeros(ns, rd)
    print(f"batched eigensolve, 256 antennas, ragged={ragged} single={single}", flush=True)
    bipp_amd.call("bipp_eigh_gram_batched", ctx.handle, ns, n, n, 0, prec, prec, wl_d.ptr,
                  s_d.ptr, n, n * n, None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, v_d.ptr, n, n * n,
                  nr_d.ptr, sc_d.ptr)
    ctx.synchronize()
    d = d_d.numpy().astype(np.float64)
    v = v_d.numpy().astype(np.complex128).transpose(0, 2, 1)  # column-major per sample
    nr = nr_d.numpy()
    tol_d, tol_v = (2e-5, 1e-4) if single else (1e-10, 1e-10)
    for k, smp in enumerate(samples):
        vo, do, nro, _ = O.eigh_gram(smp.wl, smp.s.astype(np.complex64), None, smp.xyz)
        assert nr[k] == nro
        assert np.abs(d[k] - do).max() < tol_d * np.abs(do).max()
        assert rel(vdv(v[k], d[k]), vdv(vo, do)) < tol_v
