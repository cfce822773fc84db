"""Train the learned first iterate for a workload's level-1 grid and evaluate
it: cfg4, the 3D Cartesian workload (64^3, BC ((1,1),(1,1),(-1,-1))), or cfg2,
the 2D cylindrical streamer workload (128 x 256, BC ((1,-1),(-1,-1))).

    python steamroller.jl_amd/train_guess.py --config cfg4 --samples 20000 --epochs 4 --out weights.npz
    python steamroller.jl_amd/train_guess.py --config cfg2 --samples 20000 --epochs 16 --out weights.npz

Prints the dataset and training progress, the held-out prediction error, and
cycles to 1e-8 on a tree of the workload (cfg4: a 3-level variant; cfg2: the
full tree) with the zero, exact-coarse and learned first iterates.
"""

import argparse
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4", choices=["cfg2", "cfg4"])
    ap.add_argument("--samples", type=int, default=20000)
    ap.add_argument("--eval-samples", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=2)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--widths", type=str, default="",
                    help="channels per level (default: cfg4 8,16,32,64; cfg2 8,16,32,64,64, five levels so that "
                         "the 16x8 coarsest grid sees the whole 128x256 level)")
    ap.add_argument("--out", type=str, default="")
    ap.add_argument("--init", type=str, default="", help="start from these weights")
    ap.add_argument("--no-solve-eval", action="store_true")
    ap.add_argument("--eval-only", action="store_true", help="with --init: no training")
    a = ap.parse_args()

    from __graft_entry__ import load_package
    load_package()
    import numpy as np
    import torch
    from srmg import configs, guess, solver

    log = lambda m: print(m, flush=True)  # noqa: E731
    gp = guess.grid_problem(a.config)
    # any context serves the grid solves (they take their own operator)
    small = configs.uniform_problem(gp.D, (1,) * gp.D, 8, 1, 0, tuple((1, 1) for _ in range(gp.D)))
    ctx = solver.Context(small.tree, small.conn, small.M)
    t0 = time.time()
    F, U = (None, None) if a.eval_only else guess.make_dataset(ctx, gp, a.samples, seed0=0, log=log)
    Fe, Ue = guess.make_dataset(ctx, gp, a.eval_samples, seed0=10_000_000, log=log)
    log(f"[dataset] {0 if a.eval_only else a.samples} + {a.eval_samples} samples in {time.time() - t0:.1f} s")
    widths = tuple(int(x) for x in (a.widths or {"cfg2": "8,16,32,64,64", "cfg4": "8,16,32,64"}[a.config]).split(","))
    net = solver.UNet(gp.D, gp.n, len(widths), widths, seed=0)
    if a.init:
        net = guess.load(a.init)
    else:
        net.output_scale = guess.output_scale(F, U)
        net.set_problem(gp.geometry, gp.bc, solver.LaplacianDiscretization(gp.D, 2))
    log(f"[net] {net.nparams} parameters, output scale {net.output_scale:.4e}")
    e0 = guess.evaluate(net, Fe, Ue)
    log(f"[eval] before training: median relative error {np.median(e0):.4f}")
    hist = []
    if not a.eval_only:
        t1 = time.time()
        hist = guess.train(net, F, U, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log)
        ttrain = time.time() - t1
        steps = a.epochs * (a.samples // a.batch)
        log(f"[train] {steps} steps in {ttrain:.1f} s ({1e3 * ttrain / max(1, steps):.2f} ms/step)")
        e1 = guess.evaluate(net, Fe, Ue)
        log(f"[eval] held-out median relative error {np.median(e1):.4f} "
            f"(90th percentile {np.percentile(e1, 90):.4f})")
        del F, U
        torch.cuda.empty_cache()
    else:
        e1 = e0
    fb = torch.empty_like(Fe[:1])
    torch.cuda.synchronize()
    t2 = time.time()
    for _ in range(20):
        net.forward(Fe[:1], fb, 1)
    log(f"[infer] one {'x'.join(map(str, gp.n))} grid: {1e3 * (time.time() - t2) / 20:.3f} ms per forward "
        "(host-timed, synchronous)")
    if a.out and not a.eval_only:
        guess.save(net, gp, a.out, {"train_samples": a.samples, "epochs": a.epochs,
                                    "eval_median_rel_err": float(np.median(e1)),
                                    "loss_history": np.array(hist)})
        log(f"[save] {a.out}")
    if a.no_solve_eval:
        return
    p = configs.make_cfg4(0, scale=0.35, levels=3) if a.config == "cfg4" else configs.make_cfg2(0)
    geom = solver.geometry_from_code(p.geom)
    disc = solver.LaplacianDiscretization(p.D, 2)
    for coarse, maxc in ((solver.COARSE_EXACT, 30), (solver.COARSE_GS, 200)):
        for mode, nsm in ((solver.GUESS_NONE, 0), (solver.GUESS_COARSE, 0), (solver.GUESS_LEARNED, 0),
                          (solver.GUESS_LEARNED, 4), (solver.GUESS_LEARNED, 16)):
            c = solver.Context(p.tree, p.conn, p.M)
            solver.set_guess(c, net)
            u, b, r, u1 = (c.field(solver.F64) for _ in range(4))
            b.upload(p.b)
            tt = time.time()
            n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-8, max_cycles=maxc, coarse=coarse,
                                guess=mode, guess_smooth=nsm)
            c.synchronize()
            k50 = min(50, n)
            log(f"[solve] coarse={coarse} guess={mode} smooth={nsm}: cycles {n}, first {h[0]:.3e}, "
                f"after 1 {h[min(1, n)]:.3e}, after {k50} {h[k50]:.3e}, last {h[-1]:.3e}, "
                f"fallbacks {solver.guess_fallbacks(c)}, {time.time() - tt:.3f} s")


if __name__ == "__main__":
    main()
