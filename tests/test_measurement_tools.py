"""The measurement helpers on synthetic profiler output (CPU): traffic_summary
turns FETCH_SIZE / WRITE_SIZE counter runs into bytes per launch, calibrated by
the stream kernel of the same run; trace_summary selects the dispatches inside
a marker window."""

import csv
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOOLS = os.path.join(ROOT, "steamroller.jl_amd")


def _counters(path, counter, rows):
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Counter_Name", "Kernel_Name", "Counter_Value"])
        w.writeheader()
        for name, vals in rows:
            for v in vals:  # one row per dispatch
                w.writerow({"Counter_Name": counter, "Kernel_Name": f"void srmg::{name}(float*, int)",
                            "Counter_Value": v})


def test_traffic_summary_calibrates_each_pair_by_its_stream_kernel(tmp_path):
    n = 1 << 20
    triad = "k_triad<double>"
    s9, s8 = "k_gs_sweep_colz<3, 8, float, -1, 1>", "k_gs_sweep_colz<3, 8, float, -1, 0>"
    # pair 1: FETCH_SIZE reports half of the stream's 16 n read bytes (factor 2),
    # WRITE_SIZE all of its 8 n written bytes; the sweep's counters average two
    # dispatches
    _counters(tmp_path / "f9.csv", "FETCH_SIZE", [(triad, [8 * n / 1024]), (s9, [1000.0, 1200.0])])
    _counters(tmp_path / "w9.csv", "WRITE_SIZE", [(triad, [8 * n / 1024]), (s9, [500.0, 500.0])])
    # pair 2: another calibration (factor 4 on reads, 2 on writes)
    _counters(tmp_path / "f8.csv", "FETCH_SIZE", [(triad, [4 * n / 1024]), (s8, [300.0])])
    _counters(tmp_path / "w8.csv", "WRITE_SIZE", [(triad, [4 * n / 1024]), (s8, [100.0])])
    out = tmp_path / "t.json"
    r = subprocess.run([sys.executable, os.path.join(TOOLS, "traffic_summary.py"), str(tmp_path / "f9.csv"),
                        str(tmp_path / "w9.csv"), str(n), str(out), str(tmp_path / "f8.csv"),
                        str(tmp_path / "w8.csv")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = json.load(open(out))
    assert d["calibration"]["fetch_factor"] == 2.0 and d["calibration"]["write_factor"] == 1.0
    k9, k8 = d["kernels"][s9], d["kernels"][s8]
    assert k9["read_bytes"] == 1100.0 * 1024 * 2 and k9["write_bytes"] == 500.0 * 1024
    assert k9["algorithmic_bytes"] == 3 * 4 * n
    assert k9["traffic_bytes"] == k9["read_bytes"] + k9["write_bytes"]
    assert k8["read_bytes"] == 300.0 * 1024 * 4 and k8["write_bytes"] == 100.0 * 1024 * 2
    assert set(d["kernels"]) == {s9, s8}  # the stream kernel is the calibration, not a row


def test_trace_summary_window_sums_the_dispatches_between_the_marks(tmp_path):
    trace = tmp_path / "kt.csv"
    with open(trace, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Kernel_Name", "Start_Timestamp", "End_Timestamp", "Grid_Size_X",
                                          "Workgroup_Size_X"])
        w.writeheader()
        # (name, start ns, end ns, grid, work-group size): two inside [1000, 9000]
        for name, s, e, g, wg in [("void srmg::k_a<1>(int)", 100, 900, 512, 256),
                                  ("void srmg::k_a<1>(int)", 2000, 3000, 512, 256),
                                  ("void srmg::k_b(int)", 4000, 7000, 1024, 256),
                                  ("void srmg::k_b(int)", 9500, 9900, 1024, 256)]:
            w.writerow({"Kernel_Name": name, "Start_Timestamp": s, "End_Timestamp": e, "Grid_Size_X": g,
                        "Workgroup_Size_X": wg})
    marks = tmp_path / "m.csv"
    with open(marks, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=["Message", "Start_Timestamp"])
        w.writeheader()
        w.writerow({"Message": "begin form 1", "Start_Timestamp": 1000})
        w.writerow({"Message": "end form 1", "Start_Timestamp": 9000})
    r = subprocess.run([sys.executable, os.path.join(TOOLS, "trace_summary.py"), str(trace), "--markers",
                        str(marks), "--window", "form 1"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "2 dispatches, kernel time 4.0 us, window 8.0 us"
    # rows by total time: k_b (3 us, 4 work groups), then k_a<1> (1 us, 2)
    assert lines[3] == "| k_b | 4 | 1 | 3.0 | 3.0 | 75.0 |"
    assert lines[4] == "| k_a<1> | 2 | 1 | 1.0 | 1.0 | 25.0 |"


def test_traffic_summary_all_reports_every_kernel_by_grid_size(tmp_path):
    n = 1 << 16
    rows = [("k_triad<double>", 8 * n / 1024, 4096), ("k_add<3>", 100.0, 256 * 5498), ("k_add<3>", 300.0, 256 * 5498),
            ("k_add<3>", 10.0, 256 * 1568)]
    for name, counter, scale in (("f.csv", "FETCH_SIZE", 1.0), ("w.csv", "WRITE_SIZE", 0.5)):
        with open(tmp_path / name, "w", newline="") as f:
            w = csv.DictWriter(f, fieldnames=["Counter_Name", "Kernel_Name", "Counter_Value", "Grid_Size",
                                              "Workgroup_Size"])
            w.writeheader()
            for k, v, g in rows:
                val = v if k.startswith("k_triad") and counter == "FETCH_SIZE" else v * scale
                if k.startswith("k_triad") and counter == "WRITE_SIZE":
                    val = 8 * n / 1024
                w.writerow({"Counter_Name": counter, "Kernel_Name": f"void srmg::{k}(double*)",
                            "Counter_Value": val, "Grid_Size": g, "Workgroup_Size": 256})
    out = tmp_path / "a.json"
    r = subprocess.run([sys.executable, os.path.join(TOOLS, "traffic_summary.py"), "--all", str(tmp_path / "f.csv"),
                        str(tmp_path / "w.csv"), str(n), str(out)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    d = json.load(open(out))
    got = {(k["kernel"], k["work_groups"]): k for k in d["all_kernels"]}
    # two dispatches on 5498 groups average; reads scaled by 2 (half reported)
    assert got[("k_add<3>", 5498)]["read_bytes"] == 200.0 * 1024 * 2
    assert got[("k_add<3>", 5498)]["write_bytes"] == 100.0 * 1024
    assert got[("k_add<3>", 1568)]["read_bytes"] == 10.0 * 1024 * 2
