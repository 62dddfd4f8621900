"""Benchmark: (timestep, channel) samples imaged per second on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C] [--no-cpu]

A step is one pass of the fused Bluebild chain over the configuration's whole batch of
synthetic samples (flag -> Gram -> generalized eigensolve -> energy levels -> image),
with the samples already resident in HBM.  `value` counts samples per second of the
mixed/batched form (float64 eigensolve and phase reduction, split-bf16 MFMA contraction or
float32-grid type-3 NUFFT), and is reported only when every level image is within 1e-4
relative L2 of a float64 GPU run of the same inputs and the level masks are identical.
The plain float32 per-sample form is timed beside it, and the CPU restatement of the
reference path is timed on the host.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import bipp_amd  # noqa: E402
from bipp_amd import hip, synthetic  # noqa: E402
from bipp_amd.hip import DeviceArray  # noqa: E402

PEAK_BF16_TFLOPS = 2500.0   # MI355X dense bf16 MFMA
PEAK_FP64_TFLOPS = 78.6     # MI355X fp64 (vector and matrix)
PEAK_HBM_GBS = 8000.0
HERE = os.path.dirname(os.path.abspath(__file__))


def measured_traffic(kernel: str, config: int):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc summary of this
    configuration (profiles/<kernel>_cfg<config>_traffic.json: FETCH_SIZE and WRITE_SIZE scaled
    by the copy-kernel calibration), or None when no such summary is committed."""
    path = os.path.join(HERE, "profiles", f"{kernel}_cfg{config}_traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        return json.load(fh)
IMAGER = {1: "steering", 2: "steering", 3: "steering", 4: "nufft", 5: "nufft"}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def upload(wl_: synthetic.Workload):
    s, xyz, wl = synthetic.stack_samples(wl_.samples, np.complex64)
    return DeviceArray.from_numpy(s), DeviceArray.from_numpy(xyz), DeviceArray.from_numpy(wl)


def level_intervals(ctx, dev, n_s, n, n_levels):
    """Eigenvalues of every sample (GPU, float64) -> KMeans intervals, as the reference's
    parameter estimation does before synthesis (python/bipp/parameter_estimator.py:55-105)."""
    s_d, x_d, wl_d = dev
    d_d = DeviceArray.zeros((n_s, n), np.float64)
    nr_d = DeviceArray.zeros(n_s, np.int32)
    sc_d = DeviceArray.zeros(n_s, np.float64)
    bipp_amd.call("bipp_eigh_gram_batched", ctx.handle, n_s, n, n, 0, 1, 1, wl_d.ptr, s_d.ptr, n,
                  n * n, None, 0, 0, x_d.ptr, n, 3 * n, d_d.ptr, n, None, n, n * n, nr_d.ptr,
                  sc_d.ptr)
    ctx.synchronize()
    d = d_d.numpy()
    return bipp_amd.infer_intervals(n_levels, 1.0, "log", 0.0, float("inf"), d), d


def masks(d, intervals):
    df = d.astype(np.float32)
    return np.stack([(df >= np.float32(lo)) & (df <= np.float32(hi)) & (df != 0)
                     for lo, hi in intervals])


def run_form(ctx, n, dev, n_s, lmn, intervals, compute, imager, batched, steps, warmup,
             tol=None, first=0, autotune=False, mark=None):
    """Time `steps` passes over samples [first, first + n_s) of the device arrays (with
    autotune, the NUFFT shapes are tuned in the first warmup pass).  `mark`, when given, runs
    right before and right after the timed passes (outside the timed region)."""
    s_d, x_d, wl_d = dev
    syn = bipp_amd.NufftSynthesis(ctx, n, n, intervals, lmn, compute=compute, imager=imager,
                                  batched=batched, tolerance=tol, autotune=autotune)
    s_ptr = s_d.ptr + first * n * n * 8
    x_ptr = x_d.ptr + first * 3 * n * 8
    w_ptr = wl_d.ptr + first * 8
    for _ in range(warmup):
        syn.reset()
        syn.collect_device(n_s, w_ptr, s_ptr, False, x_ptr)
    hip.synchronize()
    syn.timings(reset=True)
    if mark:
        mark()
    times = []
    for _ in range(steps):
        syn.reset()
        hip.synchronize()
        t0 = time.perf_counter()
        syn.collect_device(n_s, w_ptr, s_ptr, False, x_ptr)
        hip.synchronize()
        times.append(time.perf_counter() - t0)
    if mark:
        mark()
    return syn, float(np.mean(times)), syn.timings()


def rel_errs(a, b):
    return [float(np.linalg.norm(a[l] - b[l]) / np.linalg.norm(b[l])) for l in range(b.shape[0])]


def boundary_margin(d, intervals):
    """Smallest |d - b| / |d| over the nonzero eigenvalues d and the finite positive interval
    boundaries b: how far every eigenvalue sits from a level switch."""
    b = np.asarray(intervals, np.float64)
    b = b[np.isfinite(b)]
    b = np.unique(b[b > 0])
    dd = np.asarray(d, np.float64).ravel()
    dd = dd[dd != 0]
    if not b.size or not dd.size:
        return None
    dd = np.sort(dd)
    pos = np.searchsorted(dd, b)
    best = np.inf
    for off in (-1, 0):
        j = np.clip(pos + off, 0, dd.size - 1)
        best = min(best, float(np.min(np.abs(dd[j] - b) / np.abs(dd[j]))))
    return best


def lapack_masks(wl_, d_gpu, intervals, n_check):
    """Level masks of the GPU eigenvalues against float64 LAPACK ?hegv (scipy zhegv, the
    reference's double path) on the first n_check samples."""
    from oracle import bipp_oracle as O

    same = True
    worst = 0.0
    for k, smp in enumerate(wl_.samples[:n_check]):
        s64 = smp.s.astype(np.complex64).astype(np.complex128)
        d = O.eigh_gram(smp.wl, s64, None, smp.xyz)[1]
        same &= bool(np.array_equal(masks(d_gpu[k][None], intervals), masks(d[None], intervals)))
        worst = max(worst, float(np.max(np.abs(d_gpu[k] - d)) / np.abs(d).max()))
    return same, worst


def cpu_baseline(wl_, intervals, d_gpu, n_samples=4, grid=192):
    """CPU restatement of the reference path (float32 chegv + closed-form image, OpenMP).

    Eigensolve timed both ways on the same samples -- one threaded-LAPACK call per sample,
    and independent single-threaded problems over all cores -- and the faster counted.  The
    image is timed on a grid x grid pixel subset and scaled to the full grid.  The chegv
    eigenvalues (the reference's single-precision path) also give the level masks the GPU's
    are compared with."""
    from oracle import cpu as ocpu

    L = ocpu.lib()
    cores = L.bippc_max_threads()
    n_full = wl_.lmn.shape[0]
    sub = synthetic.make_lmn_grid(grid, math.radians(wl_.fov_deg))
    smp_eig = wl_.samples[:max(n_samples, 2 * cores)]
    s_b = np.stack([x.s for x in smp_eig])
    x_b = np.stack([x.xyz for x in smp_eig])
    w_b = np.array([x.wl for x in smp_eig])
    t0 = time.perf_counter()
    vb, db, nrb, scb = ocpu.eigh_gram_batch(w_b, s_b, x_b)
    t_batch = (time.perf_counter() - t0) / len(smp_eig)
    t_eig = t_img = 0.0
    L.bippc_lapack_threads(cores)
    for k, smp in enumerate(wl_.samples[:n_samples]):
        t0 = time.perf_counter()
        v, d, nr, sc = ocpu.eigh_gram_c(smp.wl, smp.s, smp.xyz)
        t_eig += time.perf_counter() - t0
        df = d.astype(np.float32)
        w = np.stack([sc * df * ((df >= np.float32(lo)) & (df <= np.float32(hi)))
                      for lo, hi in intervals]).astype(np.float32)
        keep = np.abs(w).sum(0) > 0
        img = np.zeros((len(intervals), sub.shape[0]), np.float32)
        t0 = time.perf_counter()
        ocpu.steering_image(v[:, keep], w[:, keep], smp.xyz / smp.wl, sub, img)
        t_img += time.perf_counter() - t0
    t_threaded = t_eig / n_samples
    eig_s = min(t_threaded, t_batch)
    scale = n_full / sub.shape[0]
    per_sample = eig_s + t_img * scale / n_samples
    m_cpu = np.stack([masks(db[k][None], intervals) for k in range(len(smp_eig))])
    m_gpu = np.stack([masks(d_gpu[k][None], intervals) for k in range(len(smp_eig))])
    return {"value": 1.0 / per_sample, "unit": "samples/s", "cores": int(cores), "kind": "port",
            "sample": (f"float32 chegv eigensolve (scipy's OpenBLAS) timed both as "
                       f"{len(smp_eig)} independent single-threaded problems over {cores} "
                       f"threads ({t_batch * 1e3:.1f} ms/sample) and as {n_samples} calls with "
                       f"{cores}-thread LAPACK ({t_threaded * 1e3:.1f} ms/sample), the faster "
                       f"counted; closed-form image (OpenMP, {cores} threads) of {n_samples} "
                       f"samples on a {grid}x{grid} pixel subset, time scaled by {scale:.1f}x to "
                       f"the full {int(math.sqrt(n_full))}^2 grid; the reference's neonufft CPU "
                       f"path cannot be built offline"),
            "eig_s_per_sample": eig_s,
            "eig_s_per_sample_threaded_lapack": t_threaded,
            "eig_s_per_sample_independent": t_batch,
            "eig_only_samples_per_s": 1.0 / eig_s,
            "image_s_per_sample_full_grid": t_img * scale / n_samples,
            "chegv_masks_identical": bool(np.array_equal(m_cpu, m_gpu)),
            "chegv_samples_compared": len(smp_eig)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # two warmup passes: the first tunes the NUFFT shapes, the second still carries one-time
    # host costs of the library's second collect (about 4 ms at configuration 4)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", type=int, default=4)
    ap.add_argument("--timesteps", type=int, default=None)
    ap.add_argument("--imager", choices=["steering", "nufft"], default=None)
    ap.add_argument("--tol", type=float, default=1e-5)
    ap.add_argument("--plain-samples", type=int, default=40)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-array (PCIe-inclusive) rate")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default NUFFT shapes (otherwise tuned during the warmup)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="end with one device-to-device copy of BYTES (hardware-counter "
                         "calibration for rocprofv3 --pmc runs)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {a.gpus}); the sample-parallel "
              f"multi-GPU form is described in DESIGN.md and not built", file=sys.stderr)
        sys.exit(2)
    imager = a.imager or IMAGER[a.config]

    t_setup = time.perf_counter()
    wl_ = synthetic.make_workload(a.config, n_timesteps=a.timesteps)
    n_s, n, P = wl_.n_samples, wl_.n_antenna, wl_.lmn.shape[0]
    log(f"config {a.config}: {n_s} samples, {n} antennas, {P} pixels, {wl_.n_levels} levels, "
        f"imager {imager} (setup {time.perf_counter() - t_setup:.1f}s)")
    ctx = bipp_amd.Context("GPU")
    dev = upload(wl_)
    intervals, d_all = level_intervals(ctx, dev, n_s, n, wl_.n_levels)
    k_eig = masks(d_all, intervals).any(axis=0).sum(axis=1)

    # mixed / batched form (the measured value)
    tune = imager == "nufft" and not a.no_autotune and a.warmup > 0

    def calibration_copy():
        """A device-to-device copy of known size on each side of the timed passes: its
        counters scale the kernels' in rocprofv3 --pmc runs, and the two copies bracket the
        dispatches of the timed passes (warmup and autotuning fall outside)."""
        src = hip.DeviceArray.zeros(a.copy_calibration, np.uint8)
        dst = hip.DeviceArray(a.copy_calibration, np.uint8)
        hip.copy_device(dst, src, a.copy_calibration)
        hip.synchronize()
        del src, dst

    syn_m, t_m, tm_m = run_form(ctx, n, dev, n_s, wl_.lmn, intervals, "mixed", imager, True,
                                a.steps, a.warmup, a.tol, autotune=tune,
                                mark=calibration_copy if a.copy_calibration > 0 else None)
    tuning = syn_m.nufft_tuning() if imager == "nufft" else None
    img_m = syn_m.images()
    d_m, _ = syn_m.last_eigenvalues()
    nb = max(tm_m["batches"], 1)
    stages = {k: v / nb for k, v in tm_m.items() if k != "batches"}
    log(f"mixed batched: {t_m * 1e3:.1f} ms/step  stages {stages}")

    # float64 GPU oracle of the same inputs: the exact steering form (full grid for the
    # steering imager, a 64x64 pixel subgrid for the NUFFT imager)
    if imager == "steering":
        qidx = np.arange(P)
    else:
        g = wl_.grid
        sub = np.arange(0, g, max(1, g // 64))
        qidx = (sub[:, None] * g + sub[None, :]).ravel()
    syn_o, t_o, _ = run_form(ctx, n, dev, n_s, wl_.lmn[qidx], intervals, "fp64", "steering",
                             True, 1, 0)
    img_o = syn_o.images()
    d_o, _ = syn_o.last_eigenvalues()
    errs_m = rel_errs(img_m[:, qidx], img_o)
    n_check = min(n_s, 16)
    masks_equal, eig_err = lapack_masks(wl_, d_m, intervals, n_check)
    margin = boundary_margin(d_m, intervals)
    log(f"fp64 oracle ({len(qidx)} pixels) {t_o * 1e3:.1f} ms; mixed rel L2 per level {errs_m}; "
        f"masks equal to LAPACK zhegv's on {n_check} samples: {masks_equal} (max eigenvalue "
        f"error {eig_err:.2e} of max |d|); min |d - boundary|/|d| over all samples {margin}")
    del syn_o

    plain = None
    if not a.no_plain:
        n_p = n_s if imager == "steering" else min(n_s, a.plain_samples)
        syn_p, t_p, _ = run_form(ctx, n, dev, n_p, wl_.lmn, intervals, "fp32", imager, False, 1,
                                 1 if imager == "steering" else 0, a.tol)
        img_p = syn_p.images()[:, qidx]
        del syn_p
        if n_p == n_s:
            ref_p = img_o
        else:
            syn_q, _, _ = run_form(ctx, n, dev, n_p, wl_.lmn[qidx], intervals, "fp64", "steering",
                                   True, 1, 0)
            ref_p = syn_q.images()
            del syn_q
        errs_p = rel_errs(img_p, ref_p)
        plain = {"samples_per_s": n_p / t_p, "ms_per_sample": t_p * 1e3 / n_p,
                 "samples_timed": n_p, "max_rel_err": max(errs_p), "per_level": errs_p,
                 "form": "float32 everywhere (phases included), one launch sequence per sample"}
        log(f"plain fp32 per-sample: {t_p * 1e3:.1f} ms for {n_p} samples, rel err {errs_p}")

    ok = max(errs_m) <= 1e-4 and masks_equal

    # the same step fed from host arrays (PCIe transfer of S and positions included)
    host_rate = None
    if not a.no_host:
        s_h, x_h, w_h = synthetic.stack_samples(wl_.samples, np.complex64)
        syn_h = bipp_amd.NufftSynthesis(ctx, n, n, intervals, wl_.lmn, compute="mixed",
                                        imager=imager, tolerance=a.tol)
        def collect_host():
            # host arrays already in the C-ABI layout (column-major per sample)
            dpp = ctypes.POINTER(ctypes.c_double)
            bipp_amd.call("bipp_synthesis_collect_host", syn_h._h, n_s, w_h.ctypes.data_as(dpp),
                          0, s_h.ctypes.data, n, n * n, None, 0, 0, x_h.ctypes.data_as(dpp), n,
                          3 * n)

        collect_host()
        hip.synchronize()
        syn_h.reset()
        t0 = time.perf_counter()
        collect_host()
        hip.synchronize()
        host_rate = n_s / (time.perf_counter() - t0)
        del syn_h
        log(f"host-array collect (PCIe included): {host_rate:.1f} samples/s")
    value = n_s / t_m if ok else 0.0

    # roofline of the dominant kernel of the timed step, plus the other stages beside it
    _, n_red = syn_m.last_eigenvalues()
    nr = n_red.astype(np.float64)
    eig_ms = stages["eigensolve_ms"]
    # eigensolver: every sample reads its complex64 pencil input (S) and antenna positions and
    # writes complex128 eigenvectors and float64 eigenvalues; 40 N^3 real flops per sample
    eig_bytes = float(np.sum(nr * nr * 8.0 + 24.0 * nr + nr * nr * 16.0 + 8.0 * nr))
    eig_flops = float(np.sum(40.0 * nr ** 3))
    eig_gbs = eig_bytes / (eig_ms * 1e-3) / 1e9
    eig_tf = eig_flops / (eig_ms * 1e-3) / 1e12
    rl_eig = {"bound": "mfma", "compute": "fp64 (vector ALU; MI355X fp64 vector and matrix "
                                         "peaks are both 78.6 TFLOP/s)",
              "achieved": eig_tf, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
              "frac": eig_tf / PEAK_FP64_TFLOPS, "traffic": None,
              "kernel": "eig_kernel (persistent batched generalized eigensolver)",
              "kernel_ms": eig_ms, "flops_model": "40 N^3 real flops per sample",
              "algorithmic_flops_per_launch": eig_flops,
              "algorithmic_bytes_per_launch": eig_bytes, "achieved_gbs": eig_gbs}
    tr = measured_traffic("eig_kernel", a.config)
    if tr:
        rl_eig["traffic"] = tr["bytes_per_launch"]
        rl_eig["traffic_source"] = tr["source"]
        rl_eig["traffic_over_algorithmic"] = tr["bytes_per_launch"] / eig_bytes
    stage_rl = {}
    if imager == "steering":
        flops = float(np.sum(8.0 * n * k_eig * P))   # complex GEMM V^T E: 8 N K P per sample
        kern_ms = stages["image_ms"]
        achieved = flops / (kern_ms * 1e-3) / 1e12
        stage_rl["steer"] = {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS,
                             "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS,
                             "kernel": "steer_kernel<BF16X3>", "kernel_ms": kern_ms,
                             "algorithmic_flops_per_launch": flops,
                             "mfma_flops_executed_per_launch": 3 * flops}
        img_ms = kern_ms  # one launch
    else:
        # type-3 NUFFT chain (sort + spread + FFT + interpolation), algorithmic bytes: every
        # point's coordinates (12 B) once and its value (8 B) per level, the spread grid
        # (nf1 cells, complex64) written by the spread, read and written by the FFT and read
        # by the interpolation per level, the pixels' lmn (12 B) and image RMW (4 B per level)
        nf = syn_m.nufft_grid()
        nf1 = [x // 2 for x in nf]
        L = len(intervals)
        m_pts = float(np.sum(nr * (nr - 1) / 2 + 1))
        g_cells = float(np.prod(nf1))
        nufft_bytes = m_pts * (12.0 + 8.0 * L) + g_cells * 8.0 * 4 * L + P * (12.0 + 8.0 * L)
        kern_ms = (stages["nufft_sort_ms"] + stages["nufft_spread_ms"] + stages["nufft_fft_ms"]
                   + stages["nufft_interp_ms"])
        achieved = nufft_bytes / (kern_ms * 1e-3) / 1e9
        stage_rl["nufft"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
                             "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
                             "kernel": "n3_points_kernel (two passes), n3_spread_kernel, "
                                       "fft_lines/fft_pass_rr (pruned), n3_interp_kernel",
                             "kernel_ms": kern_ms, "fine_grid": list(nf), "spread_grid": nf1,
                             "points": m_pts, "algorithmic_bytes_per_launch": nufft_bytes}
        # measured HBM bytes of the chain: per-dispatch counters times dispatches per step
        # (one spread launch per 4 levels; strided axes above 128 points take two passes)
        launches = {"n3_points_kernel": 2,
                    "n3_spread_kernel": (L + 3) // 4, "fft_lines_kernel": 1,
                    "fft_pass_rr_kernel": sum(2 if x > 128 else 1 for x in nf[1:]),
                    "n3_interp_kernel": 1}
        per_kernel = {}
        for k, cnt in launches.items():
            t = measured_traffic(k, a.config)
            if t:
                per_kernel[k] = {"bytes_per_dispatch": t["bytes_per_launch"], "dispatches": cnt}
        if len(per_kernel) == len(launches):
            tot = sum(v["bytes_per_dispatch"] * v["dispatches"] for v in per_kernel.values())
            stage_rl["nufft"]["traffic"] = tot
            stage_rl["nufft"]["traffic_per_kernel"] = per_kernel
            stage_rl["nufft"]["traffic_over_algorithmic"] = tot / nufft_bytes
            stage_rl["nufft"]["traffic_source"] = ("rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE "
                                                   "summaries in profiles/")
        # the chain's longest stage (its kernels run one after another)
        img_ms = max(stages[k] for k in ("nufft_sort_ms", "nufft_spread_ms", "nufft_fft_ms",
                                         "nufft_interp_ms"))
    # the dominant kernel of the step is reported as "roofline"
    if eig_ms >= img_ms or not stage_rl:
        roofline = rl_eig
    else:
        key = next(iter(stage_rl))
        roofline = stage_rl.pop(key)
        stage_rl["eigensolve"] = rl_eig
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(wl_, intervals, d_m)
            log(f"cpu baseline {cpu}")
        except Exception as exc:  # the CPU leg must not hide the GPU result
            log(f"cpu baseline failed: {exc}")
    out = {
        "metric": "timesteps*channels imaged/s at rel err <=1e-4 vs FP64 (mixed/batched)",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": t_m * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("bf16x3-mfma/fp32-acc, fp64 eigensolve" if imager == "steering"
                  else "fp32 nufft grid, fp64 eigensolve and phases"),
        "data": "synthetic (seeded disk array, point-source sky, noise 0.1)",
        "config": {"workload": wl_.name, "samples": n_s, "antennas": n, "pixels": P,
                   "levels": len(intervals), "imager": imager, "batch": n_s,
                   "nufft_tolerance": a.tol if imager == "nufft" else None},
        "accuracy": {"max_rel_l2_vs_fp64": max(errs_m), "per_level": errs_m,
                     "pixels_checked": int(len(qidx)),
                     "level_masks_equal_lapack_zhegv": masks_equal,
                     "lapack_samples_compared": n_check,
                     "max_eig_error_vs_zhegv": eig_err,
                     "level_masks_equal_chegv": cpu["chegv_masks_identical"] if cpu else None,
                     "min_boundary_margin": margin},
        "value_pcie_inclusive": host_rate,
        "plain_fp32": plain,
        "speedup_vs_plain": (value / plain["samples_per_s"]) if (plain and value) else None,
        "stages_ms_per_step": stages,
        "nufft_tuning": ({"shape": tuning[0], "candidates_timed": len(tuning[1]),
                          "fastest_ms": min((r[1] for r in tuning[1]), default=None),
                          "slowest_ms": max((r[1] for r in tuning[1]), default=None)}
                         if tuning else None),
        "roofline": roofline,
        "roofline_stages": stage_rl,
        "cpu_baseline": cpu,
    }
    if a.copy_calibration > 0:
        out["copy_calibration_bytes"] = a.copy_calibration
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
