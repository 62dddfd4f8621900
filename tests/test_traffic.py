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


def test_counter_window_between_copies(tmp_path):
    d = tmp_path / "ctr"
    d.mkdir()
    # (dispatch, kernel, SQ_WAVES, SQ_INSTS_VALU, start ns, end ns)
    rows = [(1, "spread_kernel", 10, 1e6, 0, 1000),            # warmup: excluded
            (2, "copyBuffer", 2048, 5e3, 2000, 3000),          # calibration copy
            (3, "spread_kernel", 10, 2.4e6, 4000, 1004000),    # timed, 1 ms
            (4, "copyBuffer", 4, 1e2, 1005000, 1006000),       # a small copy in the window
            (5, "copyBuffer", 2048, 5e3, 1007000, 1008000),    # calibration copy
            (6, "spread_kernel", 10, 9e9, 1009000, 1010000)]   # after: excluded
    fields = ["Dispatch_Id", "Kernel_Name", "Counter_Name", "Counter_Value", "VGPR_Count",
              "Start_Timestamp", "End_Timestamp"]
    with open(d / "run_counter_collection.csv", "w", newline="") as fh:
        w = csv.DictWriter(fh, fields)
        w.writeheader()
        for did, n, waves, valu, t0, t1 in rows:
            for c, v in (("SQ_WAVES", waves), ("SQ_INSTS_VALU", valu)):
                for part in (0.5, 0.5):   # two instances per dispatch, summed by the reader
                    w.writerow({"Dispatch_Id": did, "Kernel_Name": n, "Counter_Name": c,
                                "Counter_Value": v * part, "VGPR_Count": 32,
                                "Start_Timestamp": t0, "End_Timestamp": t1})
    out = subprocess.run([sys.executable, "-m", "bipp_amd.kernel_window", "--counters", str(d)],
                         check=True, cwd=ROOT, capture_output=True, text=True).stdout
    body = [l for l in out.splitlines() if l.startswith("| `")]
    assert len(body) == 1 and "`spread_kernel`" in body[0]
    cells = [c.strip() for c in body[0].split("|")[1:-1]]
    assert cells[1] == "1.000" and cells[2] == "32"
    assert cells[3] == "2.4e+06" and cells[4] == "10"
    # 4 cycles per wave64 instruction over 1024 SIMDs x 2.4e9 Hz x 1 ms
    assert cells[5] == f"{4 * 2.4e6 / (1024 * 2.4e9 * 1e-3):.2f}"
