"""Write the committed golden fixtures under tests/golden/.

Run from the repository root:  python tests/golden/make_golden.py [--reference DIR]

* lofar_input.npz -- the reference's own test input tests/data/lofar_input.json
  (2 LOFAR timesteps, 55 antennas, beamforming W, visibilities S, uvw pre-scaled by
  2pi/lambda, energy intervals), converted to arrays.  Needs the reference checkout.
* lofar_expected.npz -- oracle outputs for that input: eigh_gram in float32 and float64
  (eigenvalues, kept-beam count, scale, eigenvectors) and the interval images on a 32x32
  lmn grid by the direct NUFT sum (float64).  The reference's own golden images
  (lofar_nufft_output_*.json) are absent from its snapshot.
* cfg1_small.npz -- a 4-timestep cut of the 37-antenna synthetic workload with the
  oracle's eigenpairs, KMeans level intervals and 48x48 images (steering identity and
  direct sum, float64).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import bipp_oracle as O  # noqa: E402
from bipp_amd import synthetic  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def convert_lofar(ref_dir: str) -> None:
    with open(os.path.join(ref_dir, "tests", "data", "lofar_input.json")) as f:
        d = json.load(f)
    xyz, uvw, w, s = [], [], [], []
    for t in d["data"]:
        # the JSON stores every matrix column by column
        xyz.append(np.array(t["xyz"]).T)
        uvw.append(np.array(t["uvw"]).T)
        w.append((np.array(t["w_real"]) + 1j * np.array(t["w_imag"])).T)
        s.append((np.array(t["s_real"]) + 1j * np.array(t["s_imag"])).T)
    np.savez_compressed(
        os.path.join(HERE, "lofar_input.npz"),
        wl=np.float64(d["wl"]), n_antenna=d["n_antenna"], n_beam=d["n_beam"],
        n_eig_int=d["n_eig_int"], intervals=np.array(d["intervals_int"], dtype=np.float64),
        xyz=np.array(xyz), uvw=np.array(uvw), w=np.array(w), s=np.array(s))


def lofar_expected() -> None:
    z = np.load(os.path.join(HERE, "lofar_input.npz"))
    wl = float(z["wl"])
    out = {}
    lmn = synthetic.make_lmn_grid(32, math.radians(10.0))
    out["lmn"] = lmn
    for tag, single in (("f32", True), ("f64", False)):
        ds, vs, nrs, scs = [], [], [], []
        imgs = np.zeros((z["intervals"].shape[0], lmn.shape[0]))
        for k in range(z["s"].shape[0]):
            rd = np.float32 if single else np.float64
            cd = np.complex64 if single else np.complex128
            v, d, nr, sc = O.eigh_gram(rd(wl), z["s"][k].astype(cd), z["w"][k].astype(cd),
                                       z["xyz"][k].astype(rd), single=single)
            ds.append(d), vs.append(v), nrs.append(nr), scs.append(sc)
            # images from the stored (float32) eigenpairs and the fixture's scaled uvw,
            # as the reference's image_synthesis does (src/nufft_synthesis.cpp:134-165)
            d32 = d.astype(np.float32)
            for lvl, (lo, hi) in enumerate(z["intervals"]):
                dm = d32 * ((d32 >= np.float32(lo)) & (d32 <= np.float32(hi)))
                vv = O.virtual_vis(float(np.float32(sc)), dm.astype(np.float64),
                                   v.astype(np.complex64).astype(np.complex128))
                imgs[lvl] += O.nuft_sum(1.0, vv, z["uvw"][k], lmn)
        imgs /= z["s"].shape[0]
        out[f"d_{tag}"] = np.array(ds)
        out[f"v_{tag}"] = np.array(vs)
        out[f"nr_{tag}"] = np.array(nrs)
        out[f"scale_{tag}"] = np.array(scs, dtype=np.float64)
        out[f"img_{tag}"] = imgs
    np.savez_compressed(os.path.join(HERE, "lofar_expected.npz"), **out)


def cfg1_small() -> None:
    wl_ = synthetic.make_workload(1, n_timesteps=4, grid=48)
    evs = []
    vs = []
    scs = []
    for smp in wl_.samples:
        v, d, nr, sc = O.eigh_gram(smp.wl, smp.s, None, smp.xyz, single=False)
        evs.append(d), vs.append(v), scs.append(sc)
    intervals = O.infer_intervals(wl_.n_levels, 1.0, "log", 0.0, float("inf"), evs)
    img_st, _ = O.synthesize(wl_.samples, intervals, wl_.lmn, method="steering")
    img_ds, _ = O.synthesize(wl_.samples, intervals, wl_.lmn, method="direct")
    s, xyz, wl = synthetic.stack_samples(wl_.samples, np.complex128)
    np.savez_compressed(
        os.path.join(HERE, "cfg1_small.npz"), s=s, xyz=xyz, wl=wl, lmn=wl_.lmn,
        d=np.array(evs), v=np.array(vs), scale=np.array(scs, dtype=np.float64),
        intervals=intervals, img_steering=img_st, img_direct=img_ds)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reference", default="/root/reference")
    args = ap.parse_args()
    if os.path.isdir(args.reference):
        convert_lofar(args.reference)
    lofar_expected()
    cfg1_small()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
