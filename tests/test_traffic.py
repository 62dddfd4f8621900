"""The HBM-traffic reader: copy calibration and the timed-pass window (CPU, synthetic CSVs)."""
import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _write(dirname, counter, rows):
    os.makedirs(dirname, exist_ok=True)
    with open(os.path.join(dirname, "run_counter_collection.csv"), "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value"])
        w.writeheader()
        for did, name, val in rows:
            # two instances per dispatch, summed by the reader
            for part in (0.25, 0.75):
                w.writerow({"Dispatch_Id": did, "Kernel_Name": name, "Counter_Name": counter,
                            "Counter_Value": val * part})


def _run(tmp_path, rows_f, rows_w):
    f, w = str(tmp_path / "f"), str(tmp_path / "w")
    _write(f, "FETCH_SIZE", rows_f)
    _write(w, "WRITE_SIZE", rows_w)
    out = str(tmp_path / "o.json")
    subprocess.run([sys.executable, "-m", "bipp_amd.traffic", "--fetch", f, "--write", w,
                    "--kernel", "spread", "--copy-bytes", str(1 << 30), "--out", out],
                   check=True, cwd=ROOT, capture_output=True)
    return json.load(open(out))


def test_window_between_calibration_copies(tmp_path):
    gib_kib = float(1 << 20)
    rows_f = [(1, "spread_kernel", 900.0),          # warmup / autotune: excluded
              (2, "copyBuffer", 0.5 * gib_kib),      # calibration (fetch reads half)
              (3, "spread_kernel", 100.0), (4, "copyBuffer", 16.0),
              (5, "spread_kernel", 300.0),
              (6, "copyBuffer", 0.5 * gib_kib),
              (7, "spread_kernel", 5000.0)]         # after the timed passes: excluded
    rows_w = [(1, "spread_kernel", 7.0), (2, "copyBuffer", gib_kib),
              (3, "spread_kernel", 10.0), (5, "spread_kernel", 30.0),
              (6, "copyBuffer", gib_kib), (7, "spread_kernel", 70.0)]
    o = _run(tmp_path, rows_f, rows_w)
    assert o["timed_window_only"] is True
    assert o["dispatches"] == {"fetch": 2, "write": 2}
    assert abs(o["copy_calibration"]["fetch_scale"] - 0.5) < 1e-12
    assert abs(o["copy_calibration"]["write_scale"] - 1.0) < 1e-12
    # mean fetch 200 KiB / 0.5 + mean write 20 KiB / 1.0
    assert abs(o["bytes_per_launch"] - 1024.0 * (400.0 + 20.0)) < 1e-6


def test_single_copy_averages_every_dispatch(tmp_path):
    gib_kib = float(1 << 20)
    rows_f = [(1, "spread_kernel", 100.0), (2, "spread_kernel", 300.0),
              (3, "copyBuffer", 0.5 * gib_kib)]
    rows_w = [(1, "spread_kernel", 10.0), (2, "spread_kernel", 30.0),
              (3, "copyBuffer", gib_kib)]
    o = _run(tmp_path, rows_f, rows_w)
    assert o["timed_window_only"] is False
    assert abs(o["bytes_per_launch"] - 1024.0 * (400.0 + 20.0)) < 1e-6


def test_kernel_window_between_copies(tmp_path):
    d = tmp_path / "kt"
    d.mkdir()
    rows = [("copyBuffer", 0, 300), ("eig_kernel", 400, 500), ("copyBuffer", 600, 610),
            ("eig_kernel", 700, 800), ("copyBuffer", 900, 1200), ("eig_kernel", 1300, 1500)]
    with open(d / "run_kernel_trace.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, ["Kernel_Name", "Start_Timestamp", "End_Timestamp"])
        w.writeheader()
        for n, a, b in rows:
            w.writerow({"Kernel_Name": n, "Start_Timestamp": a * 1000, "End_Timestamp": b * 1000})
    out = subprocess.run([sys.executable, "-m", "bipp_amd.kernel_window", "--trace", str(d)],
                         check=True, cwd=ROOT, capture_output=True, text=True).stdout
    eig = [l for l in out.splitlines() if "`eig_kernel`" in l][0]
    # the two dispatches between the copies, 0.1 ms each; the one after them is excluded
    assert eig.split("|")[2].strip() == "2" and eig.split("|")[3].strip() == "0.100"
