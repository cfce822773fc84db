"""Summarize two rocprofv3 counter runs of kernel_bench.py (FETCH_SIZE, then
WRITE_SIZE; KiB per dispatch) into HBM bytes per launch, calibrated by the
stream kernel r = u + b of the same run (known bytes: reads 2 w N, writes w N).

    python steamroller.jl_amd/traffic_summary.py FETCH.csv WRITE.csv N_CELLS OUT.json [FETCH2.csv WRITE2.csv ...]

Writes OUT.json (per kernel: raw counters, calibration factors, traffic bytes)
and prints a text table. Further (FETCH, WRITE) pairs of runs of other tile
shapes add their sweep kernels, each pair calibrated by its own stream kernel.
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


def main():
    ncells, out = int(sys.argv[3]), sys.argv[4]
    pairs = [(sys.argv[1], sys.argv[2])] + list(zip(sys.argv[5::2], sys.argv[6::2]))
    res = {"calibration": None, "kernels": {}}
    for fetch_csv, write_csv in pairs:
        add_pair(res, fetch_csv, write_csv, ncells)
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
