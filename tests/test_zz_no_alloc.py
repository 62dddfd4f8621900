"""The no-allocation batched eigensolve (bipp_eigh_gram_batched_noalloc) against the
allocating entry point bipp_eigh_gram_batched, through the C-ABI: two calls in a row, the
second smaller, results identical (the same kernels solve each sample independently)."""
import numpy as np
import pytest

import bipp_amd
from bipp_amd import synthetic
from bipp_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu


def batch(seed, ns, n, live_lo):
    r = np.random.default_rng(seed)
    p0 = synthetic.disk_positions(r, n, 400.0)
    src, flux = synthetic.zenith_cap_sources(r, 10, 5.0)
    samples = []
    for k in range(ns):
        act = np.zeros(n, bool)
        act[r.choice(n, size=int(r.integers(live_lo, n + 1)), replace=False)] = True
        wl = 1.8 + 0.05 * k
        s = synthetic.visibilities(p0, wl, src, flux, 0.1)
        s[~act] = 0
        s[:, ~act] = 0
        samples.append(synthetic.Sample(0, k, wl, p0, s, act))
    return synthetic.stack_samples(samples, np.complex64)


def solve(ctx, entry, s_cm, x_cm, wls, compute):
    ns, n = s_cm.shape[0], s_cm.shape[1]
    dt, ct = (np.float64, np.complex128) if compute == 1 else (np.float32, np.complex64)
    s_d, x_d, wl_d = (DeviceArray.from_numpy(a) for a in (s_cm, x_cm, wls))
    d_d = DeviceArray.zeros((ns, n), dt)
    v_d = DeviceArray.zeros((ns, n, n), ct)
    nr_d = DeviceArray.zeros(ns, np.int32)
    sc_d = DeviceArray.zeros(ns, dt)
    bipp_amd.call(entry, ctx.handle, ns, n, n, 0, compute, compute, wl_d.ptr, s_d.ptr, n, n * n,
                  None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, v_d.ptr, n, n * n, nr_d.ptr,
                  sc_d.ptr)
    ctx.synchronize()
    return d_d.numpy(), v_d.numpy(), nr_d.numpy(), sc_d.numpy()


@pytest.mark.parametrize("compute", [1, 0])
def test_noalloc_equals_allocating_entry(ctx, compute):
    for seed, ns, n, lo in ((1, 12, 96, 60), (2, 5, 40, 20)):  # the second call smaller
        print(f"no-alloc eigensolve: compute {compute}, {ns} samples of {n} beams", flush=True)
        s_cm, x_cm, wls = batch(seed, ns, n, lo)
        ref = solve(ctx, "bipp_eigh_gram_batched", s_cm, x_cm, wls, compute)
        got = solve(ctx, "bipp_eigh_gram_batched_noalloc", s_cm, x_cm, wls, compute)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_noalloc_rejects_oversized_request(ctx):
    n, ns = 256, 200  # 200 MiB of reduced eigenvectors alone
    s_d = DeviceArray.zeros((ns, n, n), np.complex64)
    x_d = DeviceArray.zeros((ns, 3, n), np.float64)
    wl_d = DeviceArray.from_numpy(np.full(ns, 2.0))
    d_d = DeviceArray.zeros((ns, n), np.float64)
    with pytest.raises(bipp_amd.InvalidParameterError, match="64 MiB"):
        bipp_amd.call("bipp_eigh_gram_batched_noalloc", ctx.handle, ns, n, n, 0, 1, 1, wl_d.ptr,
                      s_d.ptr, n, n * n, None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, None, n, n * n,
                      None, None)
    with pytest.raises(bipp_amd.InvalidParameterError, match="identity beamforming"):
        bipp_amd.call("bipp_eigh_gram_batched_noalloc", ctx.handle, 1, n, n, 0, 1, 1, wl_d.ptr,
                      s_d.ptr, n, n * n, s_d.ptr, n, n * n, x_d.ptr, n, 3 * n, d_d.ptr, n, None,
                      n, n * n, None, None)
