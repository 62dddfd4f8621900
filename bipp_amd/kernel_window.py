"""Per-kernel durations of a rocprofv3 kernel trace, restricted to bench.py's timed passes.

    rocprofv3 --kernel-trace --output-format csv -d DIR -o run -- \\
        python3 bench.py ... --copy-calibration 1073741824
    python -m bipp_amd.kernel_window --trace DIR [--steps K]

bench.py --copy-calibration runs a device-to-device copy of the given size right before and
right after its timed passes; those two copies are by far the longest dispatches of the copy
kernel, and every dispatch that starts between them belongs to the timed passes (warmup,
autotuning and the accuracy checks fall outside).  Prints a markdown table: calls, average,
total and share per kernel, and per step when --steps is given.

    rocprofv3 --pmc SQ_WAVES SQ_INSTS_VALU ... --output-format csv -d DIR -o run -- \
        python3 bench.py ... --copy-calibration 1073741824
    python -m bipp_amd.kernel_window --counters DIR

prints one row per dispatch between the two calibration copies (the copies with the most
waves of the copy kernel): its duration, registers and every counter summed over its
instances, and for SQ_INSTS_VALU the share of the SIMD cycles a wave64 instruction would
occupy (4 cycles on a 16-lane SIMD, 1024 SIMDs at 2.4 GHz).
"""
from __future__ import annotations

import argparse
import csv
import glob
import os
from collections import defaultdict


def load(run_dir: str):
    files = glob.glob(os.path.join(run_dir, "**", "*kernel_trace.csv"), recursive=True)
    if not files:
        raise SystemExit(f"no kernel_trace.csv under {run_dir}")
    rows = []
    for fn in files:
        with open(fn) as fh:
            for r in csv.DictReader(fh):
                t0, t1 = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
                rows.append((fn, t0, t1, r["Kernel_Name"]))
    return rows


def window(rows, copy_kernel: str = "copyBuffer"):
    """Dispatches strictly between the first and the last calibration copy (copies lasting at
    least half as long as the longest copy) of the file that holds them."""
    copies = [r for r in rows if copy_kernel in r[3]]
    if not copies:
        raise SystemExit("no copy dispatches in the trace")
    top = max(r[2] - r[1] for r in copies)
    cal = [r for r in copies if r[2] - r[1] >= 0.5 * top]
    if len(cal) < 2:
        raise SystemExit("the trace does not hold two calibration copies")
    fn = cal[0][0]
    lo = min(r[2] for r in cal if r[0] == fn)
    hi = max(r[1] for r in cal if r[0] == fn)
    return [r for r in rows if r[0] == fn and lo <= r[1] and r[2] <= hi], (hi - lo) * 1e-6


def load_counters(run_dir: str):
    """Per dispatch of a counter_collection.csv: (dispatch id, row of its first instance,
    {counter: value summed over the instances})."""
    files = glob.glob(os.path.join(run_dir, "**", "*counter_collection.csv"), recursive=True)
    if len(files) != 1:
        raise SystemExit(f"expected one counter_collection.csv under {run_dir}, found {len(files)}")
    meta, vals = {}, defaultdict(lambda: defaultdict(float))
    with open(files[0]) as fh:
        for r in csv.DictReader(fh):
            d = int(r["Dispatch_Id"])
            meta.setdefault(d, r)
            vals[d][r["Counter_Name"]] += float(r["Counter_Value"])
    return [(d, meta[d], dict(vals[d])) for d in sorted(meta)]


def counter_window(disp, copy_kernel: str = "copyBuffer", size_counter: str = "SQ_WAVES"):
    """Dispatches strictly between the first and the last calibration copy: the copies whose
    size_counter is at least half the largest copy's."""
    copies = [x for x in disp if copy_kernel in x[1]["Kernel_Name"]]
    if not copies:
        raise SystemExit("no copy dispatches in the counter file")
    top = max(x[2].get(size_counter, 0.0) for x in copies)
    cal = [x[0] for x in copies if x[2].get(size_counter, 0.0) >= 0.5 * top]
    if len(cal) < 2:
        raise SystemExit("the counter file does not hold two calibration copies")
    return [x for x in disp if cal[0] < x[0] < cal[-1] and copy_kernel not in x[1]["Kernel_Name"]]


def counter_table(sel, simds: int = 1024, clock_hz: float = 2.4e9) -> str:
    names = sorted({c for _, _, v in sel for c in v})
    hdr = "| kernel | ms | VGPRs | " + " | ".join(names)
    sep = "|---|---:|---:|" + "---:|" * len(names)
    valu = "SQ_INSTS_VALU" in names
    if valu:
        hdr += " | VALU share"
        sep += "---:|"
    lines = [hdr + " |", sep]
    for _, m, v in sel:
        ms = (int(m["End_Timestamp"]) - int(m["Start_Timestamp"])) * 1e-6
        line = (f"| `{short(m['Kernel_Name'])}` | {ms:.3f} | {m.get('VGPR_Count', '')} | " +
                " | ".join(f"{v.get(c, 0.0):.3g}" for c in names))
        if valu:
            cycles = simds * clock_hz * ms * 1e-3
            line += f" | {4.0 * v['SQ_INSTS_VALU'] / cycles:.2f}" if cycles > 0 else " | "
        lines.append(line + " |")
    return "\n".join(lines)


def short(name: str, width: int = 70) -> str:
    name = name.replace("(anonymous namespace)::", "").split("(")[0]
    if name.startswith("void "):
        name = name[5:]
    return name if len(name) <= width else name[:width - 3] + "..."


def main() -> None:
    ap = argparse.ArgumentParser()
    src = ap.add_mutually_exclusive_group(required=True)
    src.add_argument("--trace")
    src.add_argument("--counters")
    ap.add_argument("--steps", type=int, default=0)
    ap.add_argument("--copy-kernel", default="copyBuffer")
    a = ap.parse_args()
    if a.counters:
        print(counter_table(counter_window(load_counters(a.counters), a.copy_kernel)))
        return
    sel, span_ms = window(load(a.trace), a.copy_kernel)
    per = defaultdict(list)
    for _, t0, t1, name in sel:
        per[short(name)].append((t1 - t0) * 1e-6)
    total = sum(sum(v) for v in per.values())
    print(f"window: {span_ms:.3f} ms wall between the calibration copies, "
          f"{total:.3f} ms of kernel time, {len(sel)} dispatches")
    hdr = "| kernel | calls | avg ms | total ms | % |"
    if a.steps:
        hdr += " ms / step |"
    print(hdr)
    print("|---|---:|---:|---:|---:|" + ("---:|" if a.steps else ""))
    for name, v in sorted(per.items(), key=lambda kv: -sum(kv[1])):
        line = (f"| `{name}` | {len(v)} | {sum(v) / len(v):.3f} | {sum(v):.2f} | "
                f"{100 * sum(v) / total:.1f} |")
        if a.steps:
            line += f" {sum(v) / a.steps:.3f} |"
        print(line)


if __name__ == "__main__":
    main()
