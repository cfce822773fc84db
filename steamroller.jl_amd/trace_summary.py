"""Summarize a rocprofv3 kernel trace (run_kernel_trace.csv).

    python steamroller.jl_amd/trace_summary.py TRACE.csv [--gap-us 200] [--kernels]
    python steamroller.jl_amd/trace_summary.py TRACE.csv --markers MARKERS.csv --window "mixed2 1"

Dispatches are sorted by start time and split into bursts wherever the device
idles for more than --gap-us microseconds (a host synchronization: the
solvers read each cycle's residual norm on the host). For each burst it prints
the number of dispatches, the device busy time (sum of kernel durations) and
the span (first start to last end); their difference is the time between
kernels inside the burst, i.e. what launching costs. --kernels adds the
per-kernel totals over the whole trace.

--markers/--window select the dispatches between the roctx marks
"begin WINDOW" and "end WINDOW" (bench.py --markers puts them around every
solve; rocprofv3 --marker-trace records them) and print them grouped by
kernel and grid size (work groups) as a table: launches, total and average
microseconds, share of the window's kernel time.
"""

import argparse
import csv
import re
from collections import defaultdict


def short(name):
    m = re.search(r"(k_\w+(<[^>]*>)?)", name)
    return m.group(1) if m else name.split("(")[0][:60]


def load(path):
    rows = []
    for r in csv.DictReader(open(path)):
        rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), short(r["Kernel_Name"]),
                     int(r["Grid_Size_X"]) // max(1, int(r["Workgroup_Size_X"]))))
    rows.sort()
    return rows


def bursts(rows, gap_ns):
    out, cur = [], [rows[0]]
    for a, b in zip(rows, rows[1:]):
        if b[0] - a[1] > gap_ns:
            out.append(cur)
            cur = [b]
        else:
            cur.append(b)
    out.append(cur)
    return out


def mark_time(path, msg):
    """Start timestamp of the roctx mark whose message is msg."""
    for r in csv.DictReader(open(path)):
        if msg in r.values():
            return int(r["Start_Timestamp"])
    raise SystemExit(f"no marker {msg!r} in {path}")


def window_table(rows, t0, t1):
    sel = [r for r in rows if r[0] >= t0 and r[1] <= t1]
    tot = defaultdict(lambda: [0, 0.0])
    for s, e, n, g in sel:
        tot[(n, g)][0] += 1
        tot[(n, g)][1] += (e - s) / 1e3
    busy = sum(v[1] for v in tot.values())
    print(f"{len(sel)} dispatches, kernel time {busy:.1f} us, window {(t1 - t0) / 1e3:.1f} us")
    print("| kernel | work groups | launches | total us | us per launch | share % |")
    print("|---|---|---|---|---|---|")
    for (n, g), (c, us) in sorted(tot.items(), key=lambda kv: -kv[1][1]):
        print(f"| {n} | {g} | {c} | {us:.1f} | {us / c:.1f} | {100 * us / busy:.1f} |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--gap-us", type=float, default=200.0)
    ap.add_argument("--kernels", action="store_true")
    ap.add_argument("--markers", help="rocprofv3 marker trace (run_marker_api_trace.csv)")
    ap.add_argument("--window", help="marker window name, e.g. 'mixed2 1'")
    a = ap.parse_args()
    rows = load(a.trace)
    if a.window:
        if not a.markers:
            raise SystemExit("--window needs --markers")
        window_table(rows, mark_time(a.markers, "begin " + a.window), mark_time(a.markers, "end " + a.window))
        return
    print(f"{len(rows)} dispatches")
    print("burst  dispatches  busy_ms  span_ms  idle_in_burst_%")
    for i, b in enumerate(bursts(rows, int(a.gap_us * 1e3))):
        busy = sum(e - s for s, e, _, _ in b) / 1e6
        span = (b[-1][1] - b[0][0]) / 1e6
        print(f"{i:5d}  {len(b):10d}  {busy:7.3f}  {span:7.3f}  {100 * (span - busy) / span if span else 0:6.1f}")
    if a.kernels:
        tot = defaultdict(lambda: [0, 0.0])
        for s, e, n, _ in rows:
            tot[n][0] += 1
            tot[n][1] += (e - s) / 1e3
        print("kernel  calls  total_us  avg_us")
        for n, (c, t) in sorted(tot.items(), key=lambda kv: -kv[1][1]):
            print(f"{n}  {c}  {t:.1f}  {t / c:.2f}")


if __name__ == "__main__":
    main()
