"""HBM traffic per launch of a kernel from two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE).

    python -m bipp_amd.traffic --fetch DIR --write DIR --kernel eig_kernel \\
        --copy-bytes N --command "..." --out profiles/eig_kernel_traffic.json

Each run's counter CSV holds, per dispatch, the counter in KiB.  The runs also hold one
device-to-device copy of N bytes (the runtime's copy kernel, which reads N and writes N
bytes); the counter it shows for that copy gives the scale factor applied to the kernel's.
"""
from __future__ import annotations

import argparse
import csv
import glob
import json
import os
from collections import defaultdict


def per_dispatch(run_dir: str, counter: str):
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
                key = (fn, row["Dispatch_Id"])
                vals[key] += float(row["Counter_Value"])
                names[key] = row["Kernel_Name"]
    return [(names[k], v) for k, v in vals.items()]


def mean_for(rows, needle: str):
    sel = [v for n, v in rows if needle in n]
    return (sum(sel) / len(sel), len(sel)) if sel else (None, 0)


def largest_for(rows, needle: str):
    """The calibration copy is by far the largest dispatch of the copy kernel."""
    sel = [v for n, v in rows if needle in n]
    return max(sel) if sel else None


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
    fk, nf = mean_for(f_rows, a.kernel)
    wk, nw = mean_for(w_rows, a.kernel)
    fc = largest_for(f_rows, a.copy_kernel)
    wc = largest_for(w_rows, a.copy_kernel)
    if fk is None or wk is None or not fc or not wc:
        raise SystemExit("kernel or copy dispatches missing from the counter files")
    expect_kib = a.copy_bytes / 1024.0
    f_scale, w_scale = fc / expect_kib, wc / expect_kib
    out = {
        "kernel": a.kernel,
        "dispatches": {"fetch": nf, "write": nw},
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
