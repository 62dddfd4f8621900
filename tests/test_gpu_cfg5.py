"""Configuration 5 geometry (512-antenna array in a 40 km disk, 400..512 live antennas per
timestep, 50 and 350 MHz, 6 energy levels, 4096^2 pixels over 5 deg): the 512-beam bucket of
the batched eigensolver against LAPACK, and the fused chain with the partitioned type-3 NUFFT
on the full grid, checked on a 64 x 64 pixel subgrid against the float64 oracle."""
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
def cfg5():
    # timesteps 0 and 200 (6000 s of rotation: baselines out of the z = 0 plane), first and
    # last channel
    wl_ = synthetic.make_workload(5, n_timesteps=2, channels=[0, 15], t_step=200)
    for smp in wl_.samples:
        smp.s = smp.s.astype(np.complex64).astype(np.complex128)
    return wl_


def test_eigh_gram_512_bucket(ctx, cfg5):
    """Problems of 257..512 beams (the 512 bucket) through the batched device entry point."""
    samples = cfg5.samples
    ns, n = len(samples), cfg5.n_antenna
    s_cm, x_cm, wls = synthetic.stack_samples(samples, np.complex64)
    s_d, x_d, wl_d = (DeviceArray.from_numpy(a) for a in (s_cm, x_cm, wls))
    d_d = DeviceArray.zeros((ns, n), np.float64)
    v_d = DeviceArray.zeros((ns, n, n), np.complex128)
    nr_d = DeviceArray.zeros(ns, np.int32)
    sc_d = DeviceArray.zeros(ns, np.float64)
    print("batched eigensolve, 512 bucket, live", [int(x.active.sum()) for x in samples], flush=True)
    bipp_amd.call("bipp_eigh_gram_batched", ctx.handle, ns, n, n, 0, 1, 1, wl_d.ptr, s_d.ptr, n,
                  n * n, None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, v_d.ptr, n, n * n, nr_d.ptr,
                  sc_d.ptr)
    ctx.synchronize()
    d, v, nr = d_d.numpy(), v_d.numpy().transpose(0, 2, 1), nr_d.numpy()
    for k, smp in enumerate(samples):
        vo, do, nro, sco = O.eigh_gram(smp.wl, smp.s, None, smp.xyz)
        print(f"  sample {k}: N={nro}, max |d - d_lapack| / max|d| "
              f"{np.abs(d[k] - do).max() / np.abs(do).max():.2e}", flush=True)
        assert nr[k] == nro and 400 <= nro <= 512
        assert np.abs(d[k] - do).max() < 1e-10 * np.abs(do).max()
        assert rel(vdv(v[k], d[k]), vdv(vo, do)) < 1e-10


def test_nufft_cfg5_geometry_subgrid(ctx, cfg5):
    wl_ = cfg5
    evs = [O.eigh_gram(x.wl, x.s, None, x.xyz)[1] for x in wl_.samples]
    intervals = O.infer_intervals(6, 1.0, "log", 0.0, float("inf"), evs)
    g = wl_.grid
    sub = np.arange(0, g, g // 64)
    qidx = (sub[:, None] * g + sub[None, :]).ravel()
    ref, _ = O.synthesize(wl_.samples, intervals, wl_.lmn[qidx])
    s = np.stack([x.s for x in wl_.samples]).astype(np.complex64)
    xyz = np.stack([x.xyz for x in wl_.samples])
    wl = np.array([x.wl for x in wl_.samples])
    for compute, tol, bound in (("mixed", 1e-5, 1e-4), ("fp64", 1e-6, 1e-5)):
        print(f"cfg5 nufft compute={compute}", flush=True)
        syn = bipp_amd.NufftSynthesis(ctx, 512, 512, intervals, wl_.lmn, compute=compute,
                                      imager="nufft", tolerance=tol)
        syn.collect(wl, s, xyz)
        img = syn.images()[:, qidx]
        print("  partition", syn.nufft_partition(), "grid", syn.nufft_grid(), syn.timings(),
              flush=True)
        for lvl in range(6):
            err = O.rel_l2(img[lvl], ref[lvl])
            print(f"  level {lvl}: rel L2 {err:.3e}")
            assert err <= bound
        del syn
