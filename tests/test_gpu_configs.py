"""The benchmark workloads at full size on the GPU, against the oracle.

* cfg4: the 3D Cartesian AMR Poisson problem (8^3 blocks, 5 levels, about 62k
  blocks, 28M leaf cells). The FP64 solve to 1e-8 (exact level-1 solve) gives
  the oracle's potential bit for bit and its residual history to 1e-12, also
  with fmg! as the first cycle (src/multigrid.jl:17-51); every
  mixed-precision form that the benchmark times stays within 1e-7 of the
  oracle's FP64 potential (src/multigrid.jl:60-106 restated by the oracle).
* cfg2: the 2D cylindrical streamer-like tree (6 levels, 7.7k blocks, 1.5M
  leaf cells): V-cycles with the reference's level-1 sweeps and with the
  exact level-1 solve bit for bit, the FP64 solve to 1e-8 bit for bit, the
  mixed-precision solve within 1e-7.
* cfg3: the batched photoionization forest (the three Helmholtz terms of
  src/photo.jl:45-67 on the cfg1 tree plus 64 uneven random trees, 67
  problems, 19M leaf cells) solved in one forest: every tree converges, and
  every tree's potential equals the oracle's same number of V-cycles on that
  tree alone, bit for bit (FP64), or within 1e-7 (FP32 with FP64 defect).
"""

import functools
import os
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

from oracle import Oracle
from srmg import configs, solver
from srmg.solver import F32, F64

pytestmark = pytest.mark.gpu

TOL = 1e-8


@functools.lru_cache(maxsize=None)
def _cfg(name):
    t = time.time()
    p = {"cfg2": configs.make_cfg2, "cfg4": configs.make_cfg4}[name](0)
    print(f"{name}: {p.tree.nblocks()} blocks, {p.n_leaf} leaf cells, levels {[len(l) for l in p.tree]} "
          f"(built in {time.time() - t:.1f}s)", flush=True)
    return p


def _valid(f, p):
    return f.u[(slice(f.G, f.G + p.M),) * p.D]


def _disc(p):
    return solver.HelmholtzDiscretization(p.D, 2, p.k2) if p.helm else solver.LaplacianDiscretization(p.D, 2)


def _rel(a, ref):
    return float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))


@functools.lru_cache(maxsize=None)
def _oracle_solve(name, fmg=False):
    """The oracle's FP64 solve from u = 0: exact level-1 solve, V(2, 4, 2)
    (fmg: the first cycle is fmg! with the same options), stop at
    resnorm(u) <= 1e-8 resnorm(0). Returns (cycles, history, u)."""
    p = _cfg(name)
    t = time.time()
    o = Oracle.for_problem(p)
    u, b, r, u1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    r0 = o.resnorm(r, b, u, p.s)
    o.fill_ghost_all(u)
    hist = [o.resnorm(r, b, u, p.s) / r0]
    while hist[-1] > TOL and len(hist) < 40:
        if fmg and len(hist) == 1:
            o.fmg(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, omega=1.0, coarse_exact=True)
        else:
            o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
        hist.append(o.resnorm(r, b, u, p.s) / r0)
    print(f"{name}: oracle{' (fmg first)' if fmg else ''} {len(hist) - 1} cycles to {hist[-1]:.3e} in "
          f"{time.time() - t:.1f}s ({os.environ.get('OMP_NUM_THREADS', 'all')} threads)", flush=True)
    return len(hist) - 1, np.asarray(hist), _valid(u, p).copy()


def _gpu_forms(p, forms):
    """Device solves of p from u = 0 to 1e-8: {form: (cycles, history, potential)}."""
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    e, be, re, ue = (ctx.field(F32) for _ in range(4))
    if any(f.endswith("learned") for f in forms) and p.D == 3:
        from srmg import guess
        solver.set_guess(ctx, guess.load(guess.cfg4_weights_path()))
    out = {}
    for form in forms:
        b.upload(p.b)
        u.zero()
        fmg = "_fmg" in form
        if form.startswith("fp64"):
            n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=TOL, max_cycles=40, first_fmg=fmg)
        else:
            inner = int(form[5]) if len(form) > 5 and form[5].isdigit() else 1
            g = solver.GUESS_LEARNED if form.endswith("learned") else solver.GUESS_ZERO
            n, h = solver.solve_mixed(u, b, e, be, re, ue, p.s, geom, p.bc, disc, tol=TOL, max_cycles=40,
                                      inner_cycles=inner, guess=g, first_fmg=fmg)
        host = p.b.like()
        u.download(host)
        out[form] = (n, np.asarray(h).copy(), _valid(host, p).copy())
        print(f"{p.name} {form}: {n} cycles, final {h[-1]:.3e}", flush=True)
    return out


def _check_against_oracle(name, forms):
    p = _cfg(name)
    n_o, h_o, u_o = _oracle_solve(name)
    got = _gpu_forms(p, forms)
    for form, (n, h, ug) in got.items():
        err = _rel(ug, u_o)
        print(f"{name} {form}: potential max rel diff vs the oracle's FP64 solve {err:.3e}", flush=True)
        assert h[-1] <= TOL
        if form.startswith("fp64"):
            # the same cycle sequence in the oracle: bit for bit
            n_f, h_f, u_f = _oracle_solve(name, fmg="_fmg" in form)
            assert n == n_f
            np.testing.assert_allclose(h, h_f, rtol=0, atol=1e-12)
            np.testing.assert_array_equal(ug, u_f)
        else:
            assert err <= 1e-7


def test_cfg4_full_size_solves_against_oracle():
    from srmg import guess
    forms = ["fp64", "fp64_fmg", "mixed", "mixed2", "mixed3", "mixed3_fmg"]
    if os.path.exists(guess.cfg4_weights_path()):
        forms += ["mixed2_learned", "mixed3_fmg_learned"]
    _check_against_oracle("cfg4", forms)


def test_cfg2_full_size_solves_against_oracle():
    _check_against_oracle("cfg2", ["fp64", "fp64_fmg", "mixed2", "mixed3_fmg"])


@pytest.mark.parametrize("coarse", [solver.COARSE_GS, solver.COARSE_EXACT])
def test_cfg2_full_size_vcycles_bitwise(coarse):
    p = _cfg("cfg2")
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    o = Oracle.for_problem(p)
    ou, ob, orr, ou1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    o.fill_ghost_all(ou)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    b.upload(p.b)
    solver.fill_ghost_all(u, p.bc)
    for _ in range(3):
        o.vcycle(ou, ob, orr, ou1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=coarse == solver.COARSE_EXACT)
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2, coarse=coarse)
    gu, gb = p.b.like(), p.b.like()
    u.download(gu)
    b.download(gb)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))
    np.testing.assert_array_equal(_valid(gb, p), _valid(ob, p))


@functools.lru_cache(maxsize=None)
def _cfg3():
    t = time.time()
    probs = list(configs.make_helmholtz_cfg1(0))
    _, k2s = configs.bourdon3_terms()
    probs += [configs.make_random_tree_problem(s, min_cells=1e4, max_cells=1e6, helm_k2=float(k2s[s % 3]))
              for s in range(64)]
    print(f"cfg3: {len(probs)} problems, {sum(p.n_leaf for p in probs)} leaf cells "
          f"(built in {time.time() - t:.1f}s)", flush=True)
    return probs


def _forest_solve(probs, mixed):
    geom = solver.geometry_from_code(probs[0].geom)
    ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], probs[0].M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    b.upload(solver.concat_fields([p.b for p in probs]))
    discs, ss = [_disc(p) for p in probs], [p.s for p in probs]
    if mixed:
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        n, h = solver.solve_mixed_batched(u, b, e, be, re, ue, ss, geom, probs[0].bc, discs, tol=TOL,
                                          max_cycles=40, inner_cycles=2)
    else:
        n, h = solver.solve_batched(u, b, r, u1, ss, geom, probs[0].bc, discs, tol=TOL, max_cycles=40)
    host = solver.concat_fields([p.b.like() for p in probs])
    u.download(host)
    parts = solver.split_field(host, [p.tree.nblocks() for p in probs])
    G, M = probs[0].b.G, probs[0].M
    return n, h, [a[(slice(G, G + M),) * 2] for a in parts]


@functools.lru_cache(maxsize=None)
def _cfg3_oracle(n):
    """Each cfg3 problem alone in the oracle: n exact-coarse V-cycles from 0."""
    t = time.time()
    out = []
    for p in _cfg3():
        o = Oracle.for_problem(p)
        u, b, r, u1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
        o.fill_ghost_all(u)
        for _ in range(n):
            o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
        out.append(_valid(u, p).copy())
    print(f"cfg3: oracle {n} cycles on each of {len(out)} problems in {time.time() - t:.1f}s", flush=True)
    return out


def test_cfg3_full_forest_fp64_equals_oracle_tree_by_tree():
    probs = _cfg3()
    n, h, got = _forest_solve(probs, mixed=False)
    print(f"cfg3 forest fp64: {n} cycles, worst final {h[:, -1].max():.3e}", flush=True)
    assert np.all(h[:, -1] <= TOL)
    ref = _cfg3_oracle(n)
    for k, (g, e) in enumerate(zip(got, ref)):
        if not np.array_equal(g, e):
            raise AssertionError(f"tree {k} ({probs[k].name}): max rel diff {_rel(g, e):.3e}")


def test_cfg3_full_forest_mixed_within_1e7_of_oracle():
    probs = _cfg3()
    n64, _, _ = _forest_solve(probs, mixed=False)
    n, h, got = _forest_solve(probs, mixed=True)
    print(f"cfg3 forest mixed2: {n} outer steps, worst final {h[:, -1].max():.3e}", flush=True)
    assert np.all(h[:, -1] <= TOL)
    ref = _cfg3_oracle(n64)
    errs = [_rel(g, e) for g, e in zip(got, ref)]
    print(f"cfg3 forest mixed2: worst potential rel diff vs the oracle {max(errs):.3e}", flush=True)
    assert max(errs) <= 1e-7
