"""Launch the finest-level smoother (FP64 and FP32) and the stream kernel of a
workload back to back, for timing and hardware-counter runs.

    python steamroller.jl_amd/kernel_bench.py --config cfg4 --cfg64 0 --cfg32 6 --reps 20
    python steamroller.jl_amd/kernel_bench.py --config cfg4 --cfg64 9 --cfg32 9 --level 4 --group-order 1

Prints the average launch time of each (HIP events on the solver's stream).
Under rocprofv3 --pmc the per-dispatch counters of these kernels give the HBM
traffic per launch; the stream kernel r = u + b (reads 2 w N, writes w N
bytes) calibrates the counters.
"""

import argparse
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4", choices=["cfg1", "cfg2", "cfg4"])
    ap.add_argument("--cfg64", type=int, default=0, help="sweep tile shape, FP64")
    ap.add_argument("--cfg32", type=int, default=6, help="sweep tile shape, FP32")
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--level", type=int, default=0, help="1-based level (default: the finest)")
    ap.add_argument("--group-order", type=int, default=0, choices=[0, 1],
                    help="sibling-group order of the column sweeps: 0 block order (default), 1 slabs")
    a = ap.parse_args()
    from __graft_entry__ import load_package
    load_package()
    from srmg import configs, solver
    from srmg.solver import F32, F64
    p = {"cfg1": configs.make_cfg1, "cfg2": configs.make_cfg2, "cfg4": configs.make_cfg4}[a.config](0)
    geom, disc = solver.geometry_from_code(p.geom), solver.LaplacianDiscretization(p.D, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    ctx.set_group_order(a.group_order)
    L = a.level or len(p.tree)
    nblk = ctx.level_info(L)[0]
    cells = nblk * p.M ** p.D
    s = p.s / 4 ** (L - 1)
    for dt, cfg, w in ((F64, a.cfg64, 8), (F32, a.cfg32, 4)):
        u, b, r = (ctx.field(dt) for _ in range(3))
        b.upload(p.b)
        ctx.set_sweep_config(dt, L, cfg)
        ms = ctx.time_kernel(solver.KERNEL_GS_SWEEP, L, u, b, r, s, geom, p.bc, disc, reps=a.reps)
        print(f"sweep {'fp64' if dt == F64 else 'fp32'} tile shape {cfg}, group order {a.group_order}: level {L}, {nblk} blocks, "
              f"{ms * 1e3:.2f} us per launch, {3 * w * cells / ms / 1e6:.0f} GB/s algorithmic", flush=True)
        if dt == F64:
            ms = ctx.time_kernel(solver.KERNEL_STREAM, L, u, b, r, s, geom, p.bc, disc, reps=a.reps)
            print(f"stream fp64 r = u + b: {ms * 1e3:.2f} us per launch, {3 * 8 * cells / ms / 1e6:.0f} GB/s, "
                  f"{2 * 8 * cells} bytes read, {8 * cells} written", flush=True)


if __name__ == "__main__":
    main()
