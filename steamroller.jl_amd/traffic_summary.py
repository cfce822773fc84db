"""Summarize two rocprofv3 counter runs of kernel_bench.py (FETCH_SIZE, then
WRITE_SIZE; KiB per dispatch) into HBM bytes per launch, calibrated by the
stream kernel r = u + b of the same run (known bytes: reads 2 w N, writes w N).

    python steamroller.jl_amd/traffic_summary.py FETCH.csv WRITE.csv N_CELLS OUT.json [FETCH2.csv WRITE2.csv ...]
    python steamroller.jl_amd/traffic_summary.py --all FETCH.csv WRITE.csv N_CELLS OUT.json

Writes OUT.json (per kernel: raw counters, calibration factors, traffic bytes)
and prints a text table. Further (FETCH, WRITE) pairs of runs of other tile
shapes add their sweep kernels, each pair calibrated by its own stream kernel.
--all reports every kernel of the first pair by name and grid size (work
groups), calibrated the same way, as HBM bytes read and written per launch
(for a solve profiled under the counters: bench.py runs the stream kernel).
"""

import csv
import json
import sys
from collections import defaultdict


def per_kernel(path, counter):
    vals = defaultdict(list)
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] != counter:
            continue
        name = row["Kernel_Name"].split("(")[0].replace("void ", "").replace("srmg::", "")
        vals[name].append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def per_launch_shape(path, counter):
    """Average counter per dispatch by (kernel, work groups)."""
    vals = defaultdict(list)
    for row in csv.DictReader(open(path)):
        if row["Counter_Name"] != counter:
            continue
        name = row["Kernel_Name"].split("(")[0].replace("void ", "").replace("srmg::", "")
        wg = int(row["Grid_Size"]) // max(1, int(row["Workgroup_Size"]))
        vals[(name, wg)].append(float(row["Counter_Value"]))
    return {k: sum(v) / len(v) for k, v in vals.items()}


def main():
    args = sys.argv[1:]
    every = "--all" in args
    if every:
        args.remove("--all")
    ncells, out = int(args[2]), args[3]
    pairs = [(args[0], args[1])] + list(zip(args[4::2], args[5::2]))
    res = {"calibration": None, "kernels": {}}
    for fetch_csv, write_csv in pairs:
        add_pair(res, fetch_csv, write_csv, ncells)
    if every:
        cal = res["calibration"]
        fetch, write = per_launch_shape(args[0], "FETCH_SIZE"), per_launch_shape(args[1], "WRITE_SIZE")
        rows = []
        for k in sorted(set(fetch) & set(write), key=lambda k: -(fetch[k] + write[k])):
            rd, wr = fetch[k] * 1024 * cal["fetch_factor"], write[k] * 1024 * cal["write_factor"]
            rows.append({"kernel": k[0], "work_groups": k[1], "read_bytes": rd, "write_bytes": wr})
            print(f"{k[0]:50s} {k[1]:7d} groups  read {rd / 1e6:9.2f} MB  write {wr / 1e6:9.2f} MB")
        res["all_kernels"] = rows
    with open(out, "w") as f:
        json.dump(res, f, indent=1)


def add_pair(res, fetch_csv, write_csv, ncells):
    fetch = per_kernel(fetch_csv, "FETCH_SIZE")
    write = per_kernel(write_csv, "WRITE_SIZE")
    stream = "k_triad<double>"
    if stream not in fetch or stream not in write:
        raise SystemExit("the runs hold no stream kernel to calibrate with")
    f_fetch = 2 * 8 * ncells / (fetch[stream] * 1024)
    f_write = 8 * ncells / (write[stream] * 1024)
    if res["calibration"] is None:
        res["calibration"] = {"kernel": stream, "read_bytes": 2 * 8 * ncells, "write_bytes": 8 * ncells,
                              "fetch_kib": fetch[stream], "write_kib": write[stream],
                              "fetch_factor": f_fetch, "write_factor": f_write}
    print(f"calibration ({stream}): FETCH_SIZE x {f_fetch:.3f}, WRITE_SIZE x {f_write:.3f}")
    for k in sorted(set(fetch) & set(write)):
        if not k.startswith("k_gs_sweep"):
            continue
        w = 8 if "double" in k else 4
        algo = 3 * w * ncells
        rd, wr = fetch[k] * 1024 * f_fetch, write[k] * 1024 * f_write
        res["kernels"][k] = {"fetch_kib": fetch[k], "write_kib": write[k], "read_bytes": rd, "write_bytes": wr,
                             "traffic_bytes": rd + wr, "algorithmic_bytes": algo,
                             "traffic_over_algorithmic": (rd + wr) / algo}
        print(f"{k:45s} read {rd / 1e6:8.1f} MB  write {wr / 1e6:8.1f} MB  traffic/algorithmic "
              f"{(rd + wr) / algo:.3f}")


if __name__ == "__main__":
    main()
