"""Write the committed fixtures of config 1 (tests/golden/cfg1_seed0.npz).

Contents, all produced by this build's own code (the reference ships no
fixtures):
* the cfg1 tree: per-level block coordinates and ids (the reference's
  BlockLayer pairs) and the four connectivity record lists, flattened;
* the seeded charge b (valid cells, host layout, seed 0);
* the oracle's parity-mode history: resnorm(u)/resnorm(0) after each of 50
  V-cycles with (ndown, ntop, nup) = (2, 4, 2), omega = 1, the reference's
  ntop Gauss-Seidel sweeps at level 1, from u = 0 with pre-filled ghosts
  (the caller sequence of src/streamer.jl:322-334), and the potential after
  them (valid cells).

    python tests/golden/make_fixtures.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

PATH = os.path.join(HERE, "cfg1_seed0.npz")
CYCLES = 50


def flat_records(tree, conn):
    from oracle import _flatten
    names = ["pairs", "pairs_off", "neighbor", "neighbor_off", "boundary", "boundary_off",
             "refboundary", "refboundary_off", "child", "child_off"]
    return dict(zip(names, _flatten(tree, conn)))


def oracle_history(p):
    from oracle import Oracle
    o = Oracle.for_problem(p)
    u, b, r, u1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    o.fill_ghost_all(u)
    r0 = o.resnorm(r, b, p.b.like(), p.s)
    hist = [o.resnorm(r, b, u, p.s) / r0]
    for _ in range(CYCLES):
        o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, omega=1.0)
        hist.append(o.resnorm(r, b, u, p.s) / r0)
    G, M = p.b.G, p.M
    return np.array(hist), u.u[(slice(G, G + M),) * p.D].copy()


def main():
    from __graft_entry__ import load_package
    load_package()
    from srmg import configs
    p = configs.make_cfg1(0)
    G, M = p.b.G, p.M
    hist, uv = oracle_history(p)
    np.savez_compressed(PATH, **flat_records(p.tree, p.conn), b=p.b.u[(slice(G, G + M),) * p.D].copy(),
                        history=hist, u=uv, s=np.float64(p.s), h=np.float64(p.h), M=np.int64(M),
                        cycles=np.int64(CYCLES))
    print(f"wrote {PATH}: {p.tree.nblocks()} blocks, history {hist[0]:.3e} .. {hist[-1]:.3e}")


if __name__ == "__main__":
    main()
