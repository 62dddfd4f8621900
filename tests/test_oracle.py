"""Pins the CPU oracle against the reference's own fixture, LAPACK, closed forms and
independent formulas (no GPU)."""
import math

import numpy as np
import pytest

from oracle import bipp_oracle as O
from bipp_amd import synthetic


def rng(seed=0):
    return np.random.default_rng(seed)


# ---------------------------------------------------------------- flagging / nVis
def test_flag_pattern_counts():
    s = np.ones((4, 4), dtype=np.complex64)
    idx, nvis = O.flag_beams(s)
    assert list(idx) == [0, 1, 2, 3] and nvis == 16
    s[2, :] = 0
    s[:, 2] = 0
    idx, nvis = O.flag_beams(s)
    assert list(idx) == [0, 1, 3] and nvis == 9


def test_flag_eps_threshold():
    eps = float(np.finfo(np.float32).eps)
    s = np.zeros((3, 3), dtype=np.complex64)
    s[1, 0] = math.sqrt(2 * eps)      # |s|^2 = 2 eps > eps: live
    s[2, 2] = math.sqrt(0.5 * eps)    # below threshold
    idx, nvis = O.flag_beams(s)
    assert list(idx) == [0, 1] and nvis == 2
    # eigh's test is |Re| >= eps or |Im| >= eps (src/eigensolver.cpp:62-63)
    idx2, nvis2 = O.flag_beams(s, mode="eigh")
    assert list(idx2) == [0, 1, 2] and nvis2 == 3


def test_lofar_fixture_counts(golden):
    z = golden("lofar_input.npz")
    for k, expect_nvis in enumerate((2842, 2840)):
        v, d, nr, sc = O.eigh_gram(np.float32(z["wl"]), z["s"][k].astype(np.complex64),
                                   z["w"][k].astype(np.complex64),
                                   z["xyz"][k].astype(np.float32), single=True)
        assert nr == 54
        assert sc == np.float32(1.0 / expect_nvis)
        # one beam has an all-zero weight column and is flagged out
        assert np.count_nonzero(np.abs(z["w"][k]).sum(0) == 0) == 1


def test_lofar_uvw_convention(golden):
    """uvw[i + N j] = 2pi/lambda (p_i - p_j) R with R a proper rotation (the field-centre
    frame): pins baseline order and sign of the reference's uvw."""
    z = golden("lofar_input.npz")
    wl = float(z["wl"])
    for k in range(z["xyz"].shape[0]):
        b = O.baselines_uvw(z["xyz"][k]) * (2 * np.pi / wl)
        r, *_ = np.linalg.lstsq(b, z["uvw"][k], rcond=None)
        assert np.abs(b @ r - z["uvw"][k]).max() < 1e-9 * np.abs(z["uvw"][k]).max()
        assert np.allclose(r.T @ r, np.eye(3), atol=1e-9)
        assert np.linalg.det(r) == pytest.approx(1.0, abs=1e-9)


# ---------------------------------------------------------------- Gram
def test_gram_closed_forms():
    wl = 2.0
    xyz = np.array([[0.0, 0, 0], [1.0, 0, 0], [0, 2.0, 0], [0.3, 0.4, 0.0]])
    g = O.gram_matrix(np.eye(4), xyz, wl)
    assert np.allclose(np.diag(g), 1.0)
    # |d| = lambda/2 and lambda -> sinc(1), sinc(2) = 0
    assert abs(g[0, 1]) < 1e-15 and abs(g[0, 2]) < 1e-15
    # |d| = 0.5 -> sinc(0.5) = 2/pi
    assert g[0, 3].real == pytest.approx(2 / np.pi, rel=1e-14)
    assert np.allclose(g, g.conj().T)
    assert np.allclose(O.gram_matrix(np.eye(4), xyz, wl, prefactor=4 * np.pi), 4 * np.pi * g)


def test_gram_congruence():
    r = rng(1)
    xyz = r.normal(size=(6, 3)) * 3
    w = r.normal(size=(6, 4)) + 1j * r.normal(size=(6, 4))
    g_i = O.gram_matrix(np.eye(6), xyz, 1.5)
    assert np.allclose(O.gram_matrix(w, xyz, 1.5), w.conj().T @ g_i @ w)


# ---------------------------------------------------------------- ?hegv
@pytest.mark.parametrize("single", [False, True])
def test_hegv_invariants(single):
    r = rng(2)
    n = 12
    a = r.normal(size=(n, n)) + 1j * r.normal(size=(n, n))
    a = a + a.conj().T
    b = r.normal(size=(n, n)) + 1j * r.normal(size=(n, n))
    b = b @ b.conj().T + n * np.eye(n)
    w, z = O.hegv(a, b, single)
    tol = 1e-4 if single else 1e-11
    assert np.all(np.diff(w) >= 0)
    assert np.abs(a @ z - b @ z * w).max() < tol * np.abs(a).max() * 10
    assert np.abs(z.conj().T @ b @ z - np.eye(n)).max() < tol * 10
    import scipy.linalg
    assert np.allclose(w, scipy.linalg.eigh(a, b, eigvals_only=True), rtol=tol, atol=tol)


def test_eigh_gram_flagged_equals_subproblem():
    wl_ = synthetic.make_workload(3, n_timesteps=1, n_channels=1, n_antenna=40)
    smp = wl_.samples[0]
    v, d, nr, sc = O.eigh_gram(smp.wl, smp.s, None, smp.xyz)
    act = np.nonzero(smp.active)[0]
    assert nr == act.size
    sub = smp.s[np.ix_(act, act)]
    g = O.sinc_matrix(smp.xyz[act], smp.wl)
    w, z = O.hegv(sub, g, False)
    assert np.allclose(d[:nr], w, rtol=1e-12, atol=1e-12)
    assert np.all(d[nr:] == 0) and np.all(v[~smp.active] == 0)
    assert sc == pytest.approx(1.0 / act.size ** 2)


# ---------------------------------------------------------------- virtual vis / images
def test_virtual_vis_identities():
    r = rng(3)
    n = 7
    v = r.normal(size=(n, n)) + 1j * r.normal(size=(n, n))
    d = r.normal(size=n)
    d[2] = 0
    vv = O.virtual_vis(0.5, d, v).reshape(n, n, order="F")
    assert np.allclose(vv, vv.conj().T)
    assert np.trace(vv).real == pytest.approx(0.5 * np.sum(d * np.sum(np.abs(v) ** 2, 0)))
    ph = np.exp(1j * r.uniform(0, 2 * np.pi, size=n))
    assert np.allclose(O.virtual_vis(0.5, d, v * ph), O.virtual_vis(0.5, d, v))


def test_direct_sum_equals_steering_identity():
    r = rng(4)
    n, k = 9, 5
    xyz = r.normal(size=(n, 3)) * 20
    v = r.normal(size=(n, k)) + 1j * r.normal(size=(n, k))
    wts = r.normal(size=k)
    lmn = synthetic.make_lmn_grid(12, math.radians(20))
    wl = 1.3
    img_st = O.steering_image(v, wts, xyz, wl, lmn)
    vv = (v * wts) @ v.conj().T
    uvw = O.baselines_uvw(xyz) * (2 * np.pi / wl)
    img_ds = O.nuft_sum(1.0, vv.reshape(-1, order="F"), uvw, lmn)
    assert np.allclose(img_st, img_ds, rtol=1e-11, atol=1e-11 * np.abs(img_ds).max())
    # a transposed baseline order (p_j - p_i) is the conjugate image: must differ
    img_bad = O.nuft_sum(1.0, vv.reshape(-1, order="C"), uvw, lmn)
    assert not np.allclose(img_bad, img_ds, rtol=1e-3)


def test_point_source_images_at_its_direction():
    """With a_k[i] = exp(-i k p_i.x_k) (python/bipp/statistics.py:110-113) the image of a
    single source peaks at +x_k."""
    r = rng(5)
    xyz = synthetic.disk_positions(r, 24, 200.0)
    wl = 2.0
    src = np.array([[0.01, -0.02, math.sqrt(1 - 0.01 ** 2 - 0.02 ** 2)]])
    s = synthetic.visibilities(xyz, wl, src, np.array([5.0]), 0.0)
    v, d, nr, sc = O.eigh_gram(wl, s, None, xyz)
    lmn = synthetic.make_lmn_grid(41, math.radians(6))
    img = O.steering_image(v, sc * d, xyz, wl, lmn)
    q = int(np.argmax(img))
    assert np.linalg.norm(lmn[q, :2] - src[0, :2]) < 2 * (2 * math.sin(math.radians(3)) / 40)
    assert img.min() > -1e-12


# ---------------------------------------------------------------- levels
def test_infer_intervals_known_clusters():
    d = np.array([1.0, 1.1, 0.9, 100, 110, 90, 1e4, 1.1e4, 0.0, -3.0])
    iv = O.infer_intervals(3, 1.0, "log", 0.0, float("inf"), d)
    c = sorted([np.exp(np.log([1.0, 1.1, 0.9]).mean()), np.exp(np.log([100, 110, 90.0]).mean()),
                np.exp(np.log([1e4, 1.1e4]).mean())])
    assert iv.shape == (3, 2)
    assert iv[0, 1] == float("inf") and iv[2, 0] == 0.0
    assert iv[0, 0] == pytest.approx((c[2] + c[1]) / 2) and iv[1, 1] == iv[0, 0]
    assert iv[1, 0] == pytest.approx((c[1] + c[0]) / 2) and iv[2, 1] == iv[1, 0]


def test_filters_closed_forms():
    d = np.array([4.0, -9.0, 0.0])
    assert np.array_equal(O.apply_filter("lsq", d), d)
    assert np.array_equal(O.apply_filter("std", d), [1, -1, 0])
    assert np.allclose(O.apply_filter("sqrt", d), [2, -3, 0])
    assert np.allclose(O.apply_filter("inv", d), [0.25, -1 / 9, 0])
    assert np.allclose(O.apply_filter("inv_sqrt", d), [0.5, -1 / 3, 0])
    assert np.array_equal(O.level_filter("lsq", 0, 5, d), [4, 0, 0])


def test_golden_cfg1_reproducible(golden):
    z = golden("cfg1_small.npz")
    assert np.allclose(z["img_steering"], z["img_direct"], rtol=1e-10,
                       atol=1e-12 * np.abs(z["img_direct"]).max())
    s = z["s"][0].T  # stored column-major per sample
    v, d, nr, sc = O.eigh_gram(z["wl"][0], s, None, z["xyz"][0].T)
    assert np.allclose(d, z["d"][0], rtol=1e-12, atol=1e-12)
