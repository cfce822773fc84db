"""Launch overhead of the batched photoionization solves: the forest of the
three Helmholtz terms on the cfg1 tree plus 64 uneven random trees (67
problems), solved to 1e-8 twice per form (FP64; FP32 with FP64 defect
correction, two FP32 V-cycles per correction or two in the first and three
in the later ones); the second solve of each form replays captured graphs.

    python steamroller.jl_amd/forest_bench.py [--autotune] [--markers]

Prints the host-timed milliseconds of the second solve of each form. Under
rocprofv3 --kernel-trace the kernel durations of the same solves give the
device busy time; the difference is what launching costs, i.e. the most a
single persistent kernel per smoother step could save. --autotune picks the
sweep tile shapes per level (FP64 and FP32) before the solves; --markers puts
roctx marks 'begin FORM I' / 'end FORM I' around every solve (for
rocprofv3 --marker-trace and trace_summary.py --window).
"""

import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def main():
    from __graft_entry__ import load_package
    load_package()
    from srmg import configs, solver
    from srmg.solver import F32, F64
    probs = list(configs.make_helmholtz_cfg1(0))
    _, k2s = configs.bourdon3_terms()
    probs += [configs.make_random_tree_problem(s, min_cells=1e4, max_cells=1e6, helm_k2=float(k2s[s % 3]))
              for s in range(64)]
    geom = solver.geometry_from_code(probs[0].geom)
    bc = probs[0].bc
    discs = [solver.HelmholtzDiscretization(p.D, 2, p.k2) for p in probs]
    ss = [p.s for p in probs]
    ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], probs[0].M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    e, be, re, ue = (ctx.field(F32) for _ in range(4))
    bh = solver.concat_fields([p.b for p in probs])
    if "--autotune" in sys.argv[1:]:
        b.upload(bh)
        t64 = ctx.autotune(u, b, ss[0], geom, bc, discs[0], reps=8)
        be.upload(bh)
        t32 = ctx.autotune(e, be, ss[0], geom, bc, discs[0], reps=8)
        print(f"tile shapes per level: fp64 {t64.tolist()} fp32 {t32.tolist()}", flush=True)
    print(f"forest: {len(probs)} problems, {sum(p.tree.nblocks() for p in probs)} blocks, "
          f"{sum(p.n_leaf for p in probs)} leaf cells", flush=True)
    mark = None
    if "--markers" in sys.argv[1:]:
        sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        from bench import roctx_marker
        mark = roctx_marker()
        if mark is None:
            raise SystemExit("--markers: no roctx library found")
    for form in ("fp64", "mixed2", "mixed2_then3"):
        for rep in range(2):
            b.upload(bh)
            u.zero()
            ctx.synchronize()
            if mark:
                mark(f"begin {form} {rep}")
            t = time.perf_counter()
            if form == "fp64":
                n, h = solver.solve_batched(u, b, r, u1, ss, geom, bc, discs, tol=1e-8, max_cycles=60,
                                            guess=solver.GUESS_ZERO)
            else:
                first, inner = (2, 3) if form == "mixed2_then3" else (0, 2)
                n, h = solver.solve_mixed_batched(u, b, e, be, re, ue, ss, geom, bc, discs, tol=1e-8,
                                                  max_cycles=60, inner_cycles=inner, inner_first=first,
                                                  guess=solver.GUESS_ZERO)
            ctx.synchronize()
            ms = 1e3 * (time.perf_counter() - t)
            if mark:
                mark(f"end {form} {rep}")
            print(f"{form} solve {rep + 1}: {n} cycles, {ms:.3f} ms, worst final {h[:, -1].max():.3e}", flush=True)


if __name__ == "__main__":
    main()
