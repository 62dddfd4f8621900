"""HBM traffic per launch of a kernel from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE).

    python -m bipp_amd.traffic --fetch DIR --write DIR --kernel eig_kernel \\
        --copy-bytes N --command "..." --out profiles/eig_kernel_traffic.json

Each run's counter CSV holds, per dispatch, the counter in KiB.  The runs also hold
device-to-device copies of N bytes (the runtime's copy kernel, which reads N and writes N
bytes); the counter shown for such a copy gives the scale factor applied to the kernel's.
When a run holds two or more of them (bench.py --copy-calibration puts one on each side of
its timed passes), only the kernel dispatches between the first and the last are averaged.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
from collections import defaultdict


def per_dispatch(run_dir: str, counter: str):
    """(file, dispatch id, kernel name, counter summed over its instances) per dispatch."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no counter_collection.csv under {run_dir}")
    vals = defaultdict(float)
    names = {}
    for fn in files:
        with open(fn) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                key = (fn, int(row["Dispatch_Id"]))
                vals[key] += float(row["Counter_Value"])
                names[key] = row["Kernel_Name"]
    return [(k[0], k[1], names[k], v) for k, v in sorted(vals.items())]


def calibration_copies(rows, needle: str):
    """The calibration copies: dispatches of the copy kernel within a factor of two of the
    largest one (every other copy of the program is far smaller)."""
    sel = [r for r in rows if needle in r[2]]
    if not sel:
        return []
    top = max(r[3] for r in sel)
    return [r for r in sel if r[3] >= 0.5 * top]


def window_mean(rows, needle: str, copies):
    """Mean counter of the kernel's dispatches, restricted to those between the first and the
    last calibration copy of the same file when there are two or more."""
    bounds = {}
    for fn, did, _, _ in copies:
        lo, hi = bounds.get(fn, (did, did))
        bounds[fn] = (min(lo, did), max(hi, did))
    windowed = len(copies) >= 2
    sel = [v for fn, did, n, v in rows if needle in n and
           (not windowed or (fn in bounds and bounds[fn][0] < did < bounds[fn][1]))]
    return (sum(sel) / len(sel), len(sel), windowed) if sel else (None, 0, windowed)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--kernel", required=True)
    ap.add_argument("--copy-bytes", type=float, required=True)
    ap.add_argument("--copy-kernel", default="copyBuffer")
    ap.add_argument("--command", default="")
    ap.add_argument("--out", required=True)
    a = ap.parse_args()
    f_rows = per_dispatch(a.fetch, "FETCH_SIZE")
    w_rows = per_dispatch(a.write, "WRITE_SIZE")
    f_cp = calibration_copies(f_rows, a.copy_kernel)
    w_cp = calibration_copies(w_rows, a.copy_kernel)
    fk, nf, f_win = window_mean(f_rows, a.kernel, f_cp)
    wk, nw, w_win = window_mean(w_rows, a.kernel, w_cp)
    if fk is None or wk is None or not f_cp or not w_cp:
        raise SystemExit("kernel or copy dispatches missing from the counter files")
    fc = sum(r[3] for r in f_cp) / len(f_cp)
    wc = sum(r[3] for r in w_cp) / len(w_cp)
    expect_kib = a.copy_bytes / 1024.0
    f_scale, w_scale = fc / expect_kib, wc / expect_kib
    out = {
        "kernel": a.kernel,
        "dispatches": {"fetch": nf, "write": nw},
        "timed_window_only": bool(f_win and w_win),
        "fetch_kib_per_dispatch": fk,
        "write_kib_per_dispatch": wk,
        "copy_calibration": {"bytes": a.copy_bytes, "fetch_kib": fc, "write_kib": wc,
                             "fetch_scale": f_scale, "write_scale": w_scale},
        "bytes_per_launch": 1024.0 * (fk / f_scale + wk / w_scale),
        "source": "rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE (separate runs), scaled by "
                  "the counters of a device-to-device copy of known size",
        "command": a.command,
    }
    with open(a.out, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
