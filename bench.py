#!/usr/bin/env python3
"""Benchmark: AMR Poisson solve to 1e-8 relative residual on one MI355X.

A step is one complete solve of the 3D Cartesian AMR Poisson problem (8^3-cell
blocks, 5 levels, about 31.7M stored / 27.8M leaf cells, three seeded branching
Gaussian filaments) from u = 0 to resnorm(u) <= 1e-8 resnorm(0), with the
right-hand side already resident in HBM. Two forms are timed: plain FP64
V-cycles, and FP64 defect correction around FP32 V-cycles (FP64 residual,
update and coarsest-level solve) with 1, 2 or 3 FP32 cycles per correction,
from u = 0 and from the learned level-1 first iterate (the package's trained
network), each also with fmg! (the reference's full multigrid) as the first
cycle. The headline is the fastest FP32 form. The smoother kernel's average
launch time is measured with HIP events on the solver's stream for the
roofline figure; its HBM traffic per launch comes from the committed counter
summary when that summary is for the same kernel. The CPU baseline is the
oracle (a C restatement of the reference algorithm) solving the same problem
to the same tolerance on the host's thread share, plus a one-thread sample;
every form's potential is compared with the oracle's.

Prints one JSON line.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from __graft_entry__ import load_package  # noqa: E402

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg4", choices=["cfg1", "cfg2", "cfg4"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batched", action="store_true", help="skip the batched photoionization measurement")
    ap.add_argument("--no-regrid", action="store_true", help="skip the regrid measurement")
    ap.add_argument("--no-order4", action="store_true", help="skip the order-4 measurement")
    ap.add_argument("--forms", default="", help="comma-separated solve forms to time (fp64 and fp64_fmg always)")
    ap.add_argument("--shapes", default="",
                    help="fixed sweep tile shapes instead of the autotuner's: 'fp64 list;fp32 list', "
                         "one shape per level, e.g. '0,0,9,8,9;0,0,9,8,9'")
    ap.add_argument("--markers", action="store_true",
                    help="roctx markers 'begin FORM I' / 'end FORM I' around every solve (for rocprofv3 --marker-trace)")
    return ap.parse_args()


def roctx_marker():
    """roctxMarkA of the profiler SDK's marker library, or None where it is absent."""
    import ctypes
    for name in ("librocprofiler-sdk-roctx.so.1", "librocprofiler-sdk-roctx.so", "libroctx64.so.4"):
        try:
            f = ctypes.CDLL(name).roctxMarkA
        except (OSError, AttributeError):
            continue
        f.argtypes = [ctypes.c_char_p]
        return lambda msg: f(msg.encode())
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                          "(sharding across GPUs is a design note, not implemented)"}))
        sys.exit(2)
    load_package()
    import numpy as np
    from srmg import configs, solver
    from srmg.solver import F32, F64

    t0 = time.time()
    maker = {"cfg1": configs.make_cfg1, "cfg2": configs.make_cfg2, "cfg4": configs.make_cfg4}[args.config]
    p = maker(args.seed)
    n_leaf, n_stored = p.n_leaf, p.n_stored
    log(f"[bench] {args.config}: {p.tree.nblocks()} blocks, {n_stored} stored, {n_leaf} leaf cells, "
        f"levels {[len(l) for l in p.tree]} (built in {time.time() - t0:.1f}s)")
    geom = solver.geometry_from_code(p.geom)
    disc = solver.LaplacianDiscretization(p.D, 2)

    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    e, be, re, ue = (ctx.field(F32) for _ in range(4))
    b.upload(p.b)
    bsave = ctx.field(F64).upload(p.b)  # b is mutated at covered cells by each solve
    ctx.synchronize()
    log(f"[bench] uploaded ({time.time() - t0:.1f}s)")

    t1 = time.time()
    if args.shapes:
        tuned64, tuned32 = (np.array([int(x) for x in part.split(",")], np.int32)
                            for part in args.shapes.split(";"))
        if len(tuned64) != len(p.tree) or len(tuned32) != len(p.tree):
            raise SystemExit(f"--shapes needs {len(p.tree)} shapes per precision")
        for lvl in range(1, len(p.tree) + 1):
            ctx.set_sweep_config(F64, lvl, int(tuned64[lvl - 1]))
            ctx.set_sweep_config(F32, lvl, int(tuned32[lvl - 1]))
        log(f"[bench] fixed sweep tile shapes per level: fp64 {tuned64.tolist()} fp32 {tuned32.tolist()}")
    else:
        tuned64 = ctx.autotune(u, b, p.s, geom, p.bc, disc, reps=8)
        e.zero(); be.upload(p.b)
        tuned32 = ctx.autotune(e, be, p.s, geom, p.bc, disc, reps=8)
        log(f"[bench] autotuned sweep tile shapes per level: fp64 {tuned64.tolist()} fp32 {tuned32.tolist()} "
            f"({time.time() - t1:.1f}s)")

    def reset():
        # restore the right-hand side, start from u = 0 (outside the timed region)
        b.upload(p.b)
        u.zero()

    # the learned level-1 first iterate (trained for this workload's level 1)
    net = None
    if args.config == "cfg4":
        from srmg import guess
        if os.path.exists(guess.cfg4_weights_path()):
            net = guess.load(guess.cfg4_weights_path())
            solver.set_guess(ctx, net)

    # every form starts from u = 0 (GUESS_ZERO: the initial norm and defect from b alone)
    Z = solver.GUESS_ZERO
    # (mixed, inner cycles, first iterate, fmg! first, inner cycles of the first correction)
    specs = {"fp64": (False, 1, Z, False, 0), "fp64_fmg": (False, 1, Z, True, 0), "mixed": (True, 1, Z, False, 0),
             "mixed2": (True, 2, Z, False, 0), "mixed3": (True, 3, Z, False, 0),
             "mixed2_then3": (True, 3, Z, False, 2), "mixed2_fmg": (True, 2, Z, True, 0),
             "mixed3_fmg": (True, 3, Z, True, 0)}
    if net is not None:
        specs["mixed2_learned"] = (True, 2, solver.GUESS_LEARNED, False, 0)
        specs["mixed2_then3_learned"] = (True, 3, solver.GUESS_LEARNED, False, 2)
        specs["mixed3_fmg_learned"] = (True, 3, solver.GUESS_LEARNED, True, 0)
    G, M = p.b.G, p.M

    def valid(f):
        return f.u[(slice(G, G + M),) * p.D]

    def rel_err(a, ref):
        return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))

    if args.forms:
        keep = set(args.forms.split(",")) | {"fp64", "fp64_fmg"}
        unknown = keep - set(specs)
        if unknown:
            raise SystemExit(f"unknown forms {sorted(unknown)}; known: {sorted(specs)}")
        specs = {k: v for k, v in specs.items() if k in keep}
    mark = roctx_marker() if args.markers else None
    if args.markers and mark is None:
        raise SystemExit("--markers: no roctx library found")

    forms, potentials, histories = {}, {}, {}
    for name, (mixed, inner, gmode, fmg, inner_first) in specs.items():
        times, cycles = [], []
        for it in range(args.warmup + args.steps):
            reset()
            ctx.synchronize()
            if mark:
                mark(f"begin {name} {it}")
            ts = time.perf_counter()
            if not mixed:
                n, hist = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-8, max_cycles=60, guess=gmode,
                                       first_fmg=fmg)
            else:
                n, hist = solver.solve_mixed(u, b, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-8,
                                             max_cycles=60, inner_cycles=inner, guess=gmode, first_fmg=fmg,
                                             inner_first=inner_first)
            ctx.synchronize()
            dt = time.perf_counter() - ts
            if mark:
                mark(f"end {name} {it}")
            if it >= args.warmup:
                times.append(dt)
                cycles.append(n)
            if hist[-1] > 1e-8:
                raise RuntimeError(f"{name} did not converge: {hist}")
        forms[name] = {"ms": 1e3 * float(np.mean(times)), "cycles": int(cycles[-1]),
                       "mcells_s": n_leaf / float(np.mean(times)) / 1e6, "final_rel_residual": float(hist[-1]),
                       "first_rel_residual": float(hist[0]),
                       "history": [float(f"{x:.3g}") for x in hist]}
        if gmode == solver.GUESS_LEARNED:
            forms[name]["guess_fallbacks"] = solver.guess_fallbacks(ctx)
        log(f"[bench] {name}: {forms[name]}")
        got = p.b.like()
        u.download(got)
        potentials[name] = valid(got).copy()
        histories[name] = np.asarray(hist).copy()
        if name != "fp64":
            # potential error of the form against the plain FP64 GPU solution
            forms[name]["potential_rel_err_vs_fp64"] = rel_err(potentials[name], potentials["fp64"])
    del got
    best = min((k for k in forms if not k.startswith("fp64")), key=lambda k: forms[k]["ms"])

    # roofline: the smoother pass on the finest level, FP64 and FP32
    L = len(p.tree)
    nblk, _, _ = ctx.level_info(L)
    kname = {t: sweep_kernel_name(p.D, p.M, t, int(tuned[L - 1]), ctx.level_grouped(L))
             for t, tuned in (("double", tuned64), ("float", tuned32))}
    cells = nblk * p.M ** p.D
    roof = {}
    for name, (uu, bb, rr, w) in {"fp64": (u, b, r, 8), "fp32": (e, be, re, 4)}.items():
        ms = ctx.time_kernel(solver.KERNEL_GS_SWEEP, L, uu, bb, rr, p.s / 4 ** (L - 1), geom, p.bc, disc, reps=40)
        algo = 3.0 * w * cells  # one red-black sweep: u read, b read, u write per cell
        roof[name] = {"ms": ms, "gbs": algo / (ms * 1e-3) / 1e9, "bytes": algo}
    ms = ctx.time_kernel(solver.KERNEL_STREAM, L, u, b, r, p.s, geom, p.bc, disc, reps=40)
    stream_gbs = 3 * 8 * cells / (ms * 1e-3) / 1e9
    stream16 = {}
    for name, (uu, bb, rr, w) in {"fp64": (u, b, r, 8), "fp32": (e, be, re, 4)}.items():
        ms = ctx.time_kernel(solver.KERNEL_STREAM16, L, uu, bb, rr, p.s, geom, p.bc, disc, reps=40)
        stream16[name] = round(3 * w * cells / (ms * 1e-3) / 1e9, 1)
    log(f"[bench] smoother: {roof}; stream r=u+b on the same level: {stream_gbs:.0f} GB/s; "
        f"16-byte nontemporal stream: {stream16} GB/s")

    cpu = None
    if not args.no_cpu:
        cpu, ora_hist, ora_u = cpu_baseline(p, n_leaf, args.cpu_threads, log)
        log(f"[bench] cpu: {cpu}")
        # every form against the oracle's FP64 solve of the same problem
        for name in forms:
            forms[name]["potential_rel_err_vs_oracle"] = rel_err(potentials[name], ora_u)
        h64 = histories["fp64"]
        if len(h64) == len(ora_hist):
            forms["fp64"]["history_max_abs_diff_vs_oracle"] = float(np.max(np.abs(h64 - ora_hist)))
        forms["fp64"]["oracle_cycles"] = len(ora_hist) - 1
        del ora_u
    del potentials

    batched = None if args.no_batched else batched_cfg3(log)
    order4 = None if args.no_order4 else order4_cfg1(log)
    # last: it changes the context's tree
    regrid = None if args.no_regrid else regrid_step(log, p, ctx, u, b, e)
    traffic = committed_traffic(kname["float"], args.config)
    ms = forms[best]["ms"]

    def rounded(d):
        return {k: (float(f"{v:.5g}") if isinstance(v, float) else v) for k, v in d.items()}

    out = {
        "metric": "Poisson solve to 1e-8 residual: Mcells/s, plain FP64 vs FP32+learned; HBM %",
        "value": round(forms[best]["mcells_s"], 3),
        "unit": "Mcells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 smoother+transfers, f64 residual/update/coarse solve",
        "data": "synthetic (" + ("seeded filament charge" if p.D == 3 else "seeded streamer channel and head")
                + " charge, generated tree)",
        "config": {"workload": workload_name(args.config, p, n_stored, n_leaf),
                   "blocks": p.tree.nblocks(), "leaf_cells": n_leaf, "stored_cells": n_stored,
                   "seed": args.seed, "vcycle": "(ndown, ntop, nup) = (2, 4, 2), exact FP64 level-1 solve"},
        "forms": dict({k: rounded(v) for k, v in forms.items()},
                      note_errors="potential_rel_err_vs_oracle: max-norm relative difference of the potential "
                                  "from the oracle's FP64 solve of the same problem to 1e-8 (the CPU baseline run)",
                      headline_form=best,
                      speedup_vs_fp64=round(forms["fp64"]["ms"] / forms[best]["ms"], 3),
                      speedup_vs_fp64_fmg=round(forms["fp64_fmg"]["ms"] / forms[best]["ms"], 3),
                      note="fp64: plain FP64 V-cycles (the reference's solver); mixedK: FP64 defect correction "
                           "around K FP32 V-cycles; _fmg: the first cycle (first correction) is fmg!; "
                           "_learned: first iterate from the network's level-1 prediction (level-1 smoothing, "
                           "zero fallback)"),
        "roofline": {
            "bound": "hbm",
            "kernel": f"{kname['float']} (level {L}, {nblk} blocks)",
            "achieved": round(roof["fp32"]["gbs"], 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(roof["fp32"]["gbs"] / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "fp64_kernel": {"kernel": kname["double"], "achieved": round(roof["fp64"]["gbs"], 1),
                            "frac": round(roof["fp64"]["gbs"] / HBM_PEAK_GBS, 4),
                            "ms": round(roof["fp64"]["ms"], 5)},
            "ms_per_launch": round(roof["fp32"]["ms"], 5),
            "measured_stream_gbs": round(stream_gbs, 1),
            "measured_stream16_gbs": stream16,
            "tile_shapes": {"fp64": tuned64.tolist(), "fp32": tuned32.tolist()},
            "algorithmic_bytes_per_launch": roof["fp32"]["bytes"],
        },
        "cpu_baseline": cpu,
        "batched_photoionization": batched,
        "regrid": regrid,
        "order4_cfg1": order4,
    }
    print(json.dumps(out), flush=True)


def batched_cfg3(log, ntrees=64):
    """Batched Helmholtz solves (cfg3): the three photoionization terms on the
    cfg1 tree plus `ntrees` uneven random trees (seeds 0..ntrees-1, 1e4-1e6
    cells, k^2 cycling over the three terms), solved to 1e-8 in one forest
    (every launch covers every tree) and one by one in their own contexts;
    FP64 and FP32 with FP64 residual. The forest's sweep tile shapes are
    autotuned first; the single contexts keep the default shapes (tuned on
    their few-microsecond launches they came out slower). Each timing is the
    second of two runs."""
    import numpy as np
    from srmg import configs, solver
    from srmg.solver import F32, F64
    t0 = time.time()
    probs = list(configs.make_helmholtz_cfg1(0))
    _, k2s = configs.bourdon3_terms()
    probs += [configs.make_random_tree_problem(s, min_cells=1e4, max_cells=1e6, helm_k2=float(k2s[s % 3]))
              for s in range(ntrees)]
    leaf = sum(p.n_leaf for p in probs)
    geom = solver.geometry_from_code(probs[0].geom)
    bc = probs[0].bc
    discs = [solver.HelmholtzDiscretization(p.D, 2, p.k2) for p in probs]
    ss = [p.s for p in probs]
    log(f"[bench] batched: {len(probs)} problems, {leaf} leaf cells (built in {time.time() - t0:.1f}s)")

    res = {}
    for mixed in (False, True):
        # one forest
        ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], probs[0].M)
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        bh = solver.concat_fields([p.b for p in probs])
        # sweep tile shapes per level, outside the timed solves
        if mixed:
            be.upload(bh)
            ctx.autotune(e, be, ss[0], geom, bc, discs[0], reps=8)
        else:
            b.upload(bh)
            ctx.autotune(u, b, ss[0], geom, bc, discs[0], reps=8)

        def run_forest():
            b.upload(bh)
            u.zero()
            ctx.synchronize()
            t = time.perf_counter()
            if mixed:
                n, h = solver.solve_mixed_batched(u, b, e, be, re, ue, ss, geom, bc, discs, tol=1e-8,
                                                  max_cycles=60, inner_cycles=2, guess=solver.GUESS_ZERO)
            else:
                n, h = solver.solve_batched(u, b, r, u1, ss, geom, bc, discs, tol=1e-8, max_cycles=60,
                                            guess=solver.GUESS_ZERO)
            ctx.synchronize()
            return time.perf_counter() - t, n, float(h[:, -1].max())
        run_forest()
        tf, nf, worst = run_forest()
        del ctx, u, b, r, u1, e, be, re, ue
        # one by one
        tseq = 0.0
        for p, d in zip(probs, discs):
            c1 = solver.Context(p.tree, p.conn, p.M)
            v, bb, rr, vv = (c1.field(F64) for _ in range(4))
            ee, bee, ree, uee = (c1.field(F32) for _ in range(4))

            def run_one():
                bb.upload(p.b)
                v.zero()
                c1.synchronize()
                t = time.perf_counter()
                if mixed:
                    solver.solve_mixed(v, bb, ee, bee, ree, uee, p.s, geom, bc, d, tol=1e-8, max_cycles=60,
                                       inner_cycles=2, guess=solver.GUESS_ZERO)
                else:
                    solver.solve(v, bb, rr, vv, p.s, geom, bc, d, tol=1e-8, max_cycles=60, guess=solver.GUESS_ZERO)
                c1.synchronize()
                return time.perf_counter() - t
            run_one()
            tseq += run_one()
        key = "fp32_mixed2" if mixed else "fp64"
        res[key] = {"forest_ms": round(1e3 * tf, 3), "one_by_one_ms": round(1e3 * tseq, 3),
                    "speedup": round(tseq / tf, 2), "forest_mcells_s": round(leaf / tf / 1e6, 1),
                    "cycles": int(nf), "worst_final_rel_residual": float(f"{worst:.4g}")}
        log(f"[bench] batched {key}: {res[key]}")
    res["workload"] = (f"cfg3: 3 photoionization Helmholtz terms on the cfg1 tree + {ntrees} random trees, "
                       f"{len(probs)} problems, {leaf} leaf cells, to 1e-8 each")
    return res


def order4_cfg1(log, cycles=10):
    """V-cycles (2, 4, 2) of the order-2 and the compact order-4 Poisson
    operator on the cfg1 tree (2D cylindrical AMR, FP64), from u = 0 with the
    same charge: milliseconds per cycle and the relative residual reached
    (resnorm! of each operator). Each order-4 cycle is replayed from its
    captured graph: the colour passes (the passes after a sweep's first read
    the copy and boundary ghosts live), one copy-and-boundary frame refill
    per sweep and the coarse-fine frame fills, with its explicit corner and
    edge ghosts."""
    from srmg import configs, solver
    from srmg.solver import F64
    p = configs.make_cfg1(0)
    geom = solver.geometry_from_code(p.geom)
    res = {}
    for order in (2, 4):
        disc = solver.LaplacianDiscretization(p.D, order)
        ctx = solver.Context(p.tree, p.conn, p.M)
        u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
        q.upload(p.b)
        solver.rhs(q1, q, geom, p.bc, disc)
        solver.fill_ghost_all(u, p.bc, disc)
        r0 = solver.resnorm(r, q1, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc)
        solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)  # warm-up
        ctx.synchronize()
        t = time.perf_counter()
        for _ in range(cycles):
            solver.fill_ghost_all(u, p.bc, disc)
            solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
        ctx.synchronize()
        ms = 1e3 * (time.perf_counter() - t) / cycles
        rn = solver.resnorm(r, q1, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc)
        res[f"order{order}"] = {"ms_per_cycle": round(ms, 3), "rel_residual_after": float(f"{rn / r0:.4g}"),
                                "cycles": cycles + 1}
    res["workload"] = (f"cfg1 tree ({p.tree.nblocks()} blocks of {p.M}^2, levels {[len(l) for l in p.tree]}), "
                       "cylindrical, FP64, V(2,4,2) with the reference's level-1 sweeps")
    log(f"[bench] order 4: {res}")
    return res


def regrid_step(log, p, ctx, u, b, e):
    """One refine! step of the workload's tree with the fields on the device:
    every eighth sibling group of the finest level removed, every eighth
    level-(L-1) leaf whose star neighbours all exist refined; the host edits the
    tree (srmg.tree.applydelta), srmg_connectivity computes connectivity! on the
    GPU (checked against the host restatement), srmg_regrid rebuilds the tables
    and carries u and b (FP64) and e (FP32) across (interp! of new blocks,
    restrict! of removed ones). The table build alone is timed on a second
    context."""
    import copy

    import numpy as np
    from srmg import solver
    from srmg.solver import REGRID_INTERP, REGRID_RESTRICT
    from srmg.tree import REFINE, REMOVE, applydelta, connectivity, parentcoord, star_stencil, subblocks
    tree, D = p.tree, p.D
    L = len(tree)
    if L < 3:
        return None
    t0 = time.time()
    delta = {}
    fin, lay = tree[L - 1], tree[L - 2]
    for k, pc in enumerate(sorted({parentcoord(c) for c, _ in fin.pairs})):
        if k % 8 == 0:
            for ch in subblocks(pc):
                delta[fin.get(ch)] = REMOVE
    cand = []
    for c, blk in lay.pairs:
        if not tree.isleaf(L - 1, c):
            continue
        nb = [tuple(x + y for x, y in zip(c, s)) for s in star_stencil(D)]
        if all(lay.hasblock(q) or not lay.in_domain(q) for q in nb):
            cand.append(blk)
    for blk in cand[::8]:
        delta[blk] = REFINE
    t2 = copy.deepcopy(tree)
    nb2, actions = applydelta(t2, delta, [], tree.nblocks())
    edit_s = time.time() - t0
    t0 = time.time()
    c_host = connectivity(t2)
    host_conn_s = time.time() - t0
    solver.connectivity_device(t2)  # first call: loads its kernels
    t0 = time.perf_counter()
    c2 = solver.connectivity_device(t2)
    dev_conn_ms = 1e3 * (time.perf_counter() - t0)
    same = all(np.array_equal(getattr(c2, k)[l].reshape(-1), getattr(c_host, k)[l].reshape(-1))
               for k in ("neighbor", "boundary", "refboundary", "child") for l in range(len(t2)))
    nnew = sum(1 for a in actions if a[0] == "interp")
    nrem = sum(1 for a in actions if a[0] == "restrict")
    both = REGRID_INTERP | REGRID_RESTRICT
    ctx.synchronize()
    ctx.regrid(t2, c2, fields=[(u, both), (b, both), (e, both)], bc=p.bc, nblocks=nb2)
    regrid_ms = ctx.last_regrid_ms
    other = solver.Context(tree, p.conn, p.M)
    other.synchronize()
    other.regrid(t2, c2, nblocks=nb2)
    tables_ms = other.last_regrid_ms
    del other
    res = {"blocks_before": tree.nblocks(), "blocks_after": t2.nblocks(), "new_blocks": nnew,
           "removed_blocks": nrem, "fields": "u, b (FP64), e (FP32): interp! + restrict!",
           "regrid_ms": round(regrid_ms, 2), "table_build_ms": round(tables_ms, 2),
           "device_connectivity_ms": round(dev_conn_ms, 2), "device_connectivity_equals_host": bool(same),
           "host_tree_edit_s": round(edit_s, 2), "host_connectivity_s": round(host_conn_s, 2),
           "note": "regrid_ms: the srmg_regrid call (host table build + device transfer of the three fields); "
                   "table_build_ms: the srmg_set_tree call for the same tree on a context without fields; "
                   "device_connectivity_ms: connectivity! by srmg_connectivity (records back on the host), "
                   "host_connectivity_s: the Python restatement; the tree edit runs in Python"}
    log(f"[bench] regrid: {res}")
    return res


def sweep_kernel_name(D, M, t, cfg, grouped):
    """Name of the sweep kernel a tile shape launches (as the profiler lists it)."""
    if cfg in (7, 8, 9, 10, 11) and grouped and (D, M) == (3, 8):
        var = {8: ", 0", 9: ", 1", 10: ", 2", 11: ", 3"}.get(cfg, "")
        return f"k_gs_sweep_col{'' if cfg == 7 else 'z'}<{D}, {M}, {t}, -1{var}>"
    if cfg >= 4 and grouped and ((D, M) in ((3, 8), (2, 16))):
        nt = {4: 256, 5: 512, 6: 1024, 7: 256, 8: 512, 9: 512, 10: 512, 11: 512}[cfg]
        if ((1 << D) * M ** D // 2) % nt == 0:
            return f"k_gs_sweep_group<{D}, {M}, {t}, {nt}, -1>"
    small = (M ** D // 2) <= 256
    bpw0 = 256 // (M ** D // 2) if small else 1
    k = cfg if cfg < 4 else 0
    bpw = bpw0 * (1 if k == 0 else (4 if k == 3 else 2)) if small else 1
    nt = (256 if k < 2 else 512) if small else (256 if k == 0 else (1024 if k == 2 else 512))
    return f"k_gs_sweep<{D}, {M}, {t}, {bpw}, {nt}, -1>"


def committed_traffic(kernel, config):
    """HBM bytes per launch of `kernel` from the committed counter summary
    (profiles/), or None when the summary is for another kernel or workload."""
    path = os.path.join(ROOT, "profiles", "sweep_traffic_cfg4.json")
    if config != "cfg4" or not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    k = d.get("kernels", {}).get(kernel)
    return None if k is None else k.get("traffic_bytes")


def workload_name(cfg, p, n_stored, n_leaf):
    geom = {0: "Cartesian", 1: "cylindrical (r, z)", 2: "cylindrical (z, r)"}[p.geom]
    return (f"{cfg}: {p.D}D {geom} AMR Poisson, {p.M}^{p.D} blocks, {len(p.tree)} levels, "
            f"{n_stored} stored / {n_leaf} leaf cells, to 1e-8 relative residual")


def cpu_model():
    """The host CPU's model name (lscpu, else /proc/cpuinfo)."""
    import subprocess
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                return line.split(":", 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(p, n_leaf, threads, log):
    """The oracle (C restatement of the reference algorithm, OpenMP over the
    blocks of a level as the reference's @batch) on the same tree, options and
    tolerance: the full FP64 solve from u = 0 to resnorm <= 1e-8 resnorm(0),
    timed, with the host's thread share (OMP_NUM_THREADS, else every CPU);
    then one V-cycle and one resnorm on one thread, timed, as the sample of the
    one-thread solve. Returns (summary, oracle history, oracle potential)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as ora
    nthreads = threads or int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
    ora.set_threads(nthreads)
    o = ora.Oracle.for_problem(p)
    u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
    t0 = time.perf_counter()
    r0 = o.resnorm(r, b, u, p.s)
    o.fill_ghost_all(u)
    hist = [o.resnorm(r, b, u, p.s) / r0]
    while hist[-1] > 1e-8 and len(hist) < 60:
        o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
        hist.append(o.resnorm(r, b, u, p.s) / r0)
    t_solve = time.perf_counter() - t0
    cycles = len(hist) - 1
    G, M = p.b.G, p.M
    pot = u.u[(slice(G, G + M),) * p.D].copy()
    log(f"[bench] cpu: oracle solve {cycles} cycles in {t_solve:.2f}s on {nthreads} threads")
    # one thread: one V-cycle and one norm (continuing from the solution)
    ora.set_threads(1)
    t0 = time.perf_counter()
    o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
    t1_cycle = time.perf_counter() - t0
    t0 = time.perf_counter()
    o.resnorm(r, b, u, p.s)
    t1_norm = time.perf_counter() - t0
    ora.set_threads(nthreads)
    t1_solve = cycles * t1_cycle + (cycles + 2) * t1_norm
    del u, r, u1, b, o
    summary = {"value": round(n_leaf / t_solve / 1e6, 4), "unit": "Mcells/s", "cores": nthreads, "kind": "port",
               "sample": f"the full FP64 solve to 1e-8 on the whole tree: {cycles} V-cycles and {cycles + 2} "
                         f"norms in {t_solve:.2f}s on {nthreads} threads",
               "cpu_model": cpu_model(), "cpu_count": os.cpu_count(),
               "final_rel_residual": float(hist[-1]),
               "one_thread": {"value": round(n_leaf / t1_solve / 1e6, 4), "unit": "Mcells/s", "cores": 1,
                              "sample": f"one V-cycle ({t1_cycle:.2f}s) and one norm ({t1_norm:.2f}s) on one "
                                        f"thread, scaled to {cycles} cycles and {cycles + 2} norms"}}
    return summary, np.asarray(hist), pot


if __name__ == "__main__":
    main()
