"""Batched launches (several independent problems per launch) against the same
problems solved one by one.

A forest holds the three photoionization Helmholtz terms on the cfg1 tree
(src/photo.jl:45-67: same source, own k^2 each, s = h^2) and uneven random
trees. Every tree must get exactly its own V-cycle: the batched results equal
the single-problem results bit for bit, with the reference's coarsest-level
treatment and with the exact coarsest-level solve alike.
"""

import numpy as np
import pytest

from srmg import configs, solver
from srmg.solver import F32, F64

pytestmark = pytest.mark.gpu


def _problems(n_random=4, max_cells=4e4):
    probs = list(configs.make_helmholtz_cfg1(0))
    _, k2s = configs.bourdon3_terms()
    for i in range(n_random):
        probs.append(configs.make_random_tree_problem(100 + i, min_cells=1e4, max_cells=max_cells,
                                                      helm_k2=float(k2s[i % 3])))
    return probs


def _disc(p):
    return solver.HelmholtzDiscretization(p.D, 2, p.k2) if p.helm else solver.LaplacianDiscretization(p.D, 2)


def _valid(a, p):
    G, M = p.b.G, p.M
    return a[(slice(G, G + M),) * p.D]


def _forest(probs, dtype=F64):
    ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], probs[0].M)
    fields = [ctx.field(dtype) for _ in range(4)]
    fields[1].upload(solver.concat_fields([p.b for p in probs]))
    return ctx, fields


def _download(f, probs):
    host = solver.concat_fields([p.b.like() for p in probs])
    f.download(host)
    return solver.split_field(host, [p.tree.nblocks() for p in probs])


@pytest.mark.parametrize("coarse", [solver.COARSE_GS, solver.COARSE_EXACT])
def test_forest_vcycles_equal_single_problem_vcycles(coarse):
    probs = _problems()
    print(f"{len(probs)} trees, coarse={coarse}", flush=True)
    geom = solver.geometry_from_code(probs[0].geom)
    ctx, (u, b, r, u1) = _forest(probs)
    solver.fill_ghost_all(u, probs[0].bc)
    for _ in range(3):
        solver.vcycle_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc,
                              [_disc(p) for p in probs], ndown=2, ntop=4, nup=2, coarse=coarse)
    got = _download(u, probs)
    gotb = _download(b, probs)
    for k, p in enumerate(probs):
        c1 = solver.Context(p.tree, p.conn, p.M)
        v, bb, rr, vv = (c1.field(F64) for _ in range(4))
        bb.upload(p.b)
        solver.fill_ghost_all(v, p.bc)
        for _ in range(3):
            solver.vcycle(v, bb, rr, vv, p.s, p.tree, p.conn, geom, p.bc, _disc(p), ndown=2, ntop=4, nup=2,
                          coarse=coarse)
        ref, refb = p.b.like(), p.b.like()
        v.download(ref)
        bb.download(refb)
        np.testing.assert_array_equal(_valid(got[k], p), _valid(ref.u, p))
        np.testing.assert_array_equal(_valid(gotb[k], p), _valid(refb.u, p))


def test_forest_solve_matches_single_solves():
    probs = _problems()
    geom = solver.geometry_from_code(probs[0].geom)
    ctx, (u, b, r, u1) = _forest(probs)
    n, hist = solver.solve_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc,
                                   [_disc(p) for p in probs], tol=1e-9)
    assert np.all(hist[:, -1] <= 1e-9)
    got = _download(u, probs)
    for k, p in enumerate(probs):
        c1 = solver.Context(p.tree, p.conn, p.M)
        v, bb, rr, vv = (c1.field(F64) for _ in range(4))
        bb.upload(p.b)
        n1, h1 = solver.solve(v, bb, rr, vv, p.s, geom, p.bc, _disc(p), tol=1e-9, max_cycles=n)
        # the batch runs until its slowest tree converges: compare the first n1
        # cycles (the per-tree norm is summed in another fixed order than a
        # lone tree's, so the histories agree to rounding)
        np.testing.assert_allclose(hist[k, : n1 + 1], h1, rtol=1e-12, atol=0)
        if n1 == n:
            ref = p.b.like()
            v.download(ref)
            np.testing.assert_array_equal(_valid(got[k], p), _valid(ref.u, p))


def test_forest_mixed_precision_matches_fp64():
    probs = _problems(n_random=3)
    geom = solver.geometry_from_code(probs[0].geom)
    discs = [_disc(p) for p in probs]
    ctx, (u, b, r, u1) = _forest(probs)
    solver.solve_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc, discs, tol=1e-8)
    ref = _download(u, probs)
    ctx2, fs = _forest(probs)
    v, b2 = fs[0], fs[1]
    e, be, re, ue = (ctx2.field(F32) for _ in range(4))
    n, hist = solver.solve_mixed_batched(v, b2, e, be, re, ue, [p.s for p in probs], geom, probs[0].bc, discs,
                                         tol=1e-8)
    assert np.all(hist[:, -1] <= 1e-8)
    got = _download(v, probs)
    for k, p in enumerate(probs):
        a, e_ = _valid(got[k], p), _valid(ref[k], p)
        err = np.max(np.abs(a - e_)) / np.max(np.abs(e_))
        print(f"tree {k}: mixed vs fp64 potential rel err {err:.2e}")
        assert err <= 1e-7


def test_single_problem_calls_on_a_forest_apply_to_every_tree():
    probs = [configs.make_random_tree_problem(200 + i, min_cells=1e4, max_cells=3e4) for i in range(3)]
    geom = solver.geometry_from_code(probs[0].geom)
    ctx, (u, b, r, u1) = _forest(probs)
    solver.fill_ghost_all(u, probs[0].bc)
    solver.vcycle(u, b, r, u1, 1.0e-3, None, None, geom, probs[0].bc, solver.LaplacianDiscretization(2, 2),
                  ndown=1, ntop=2, nup=1)
    got = _download(u, probs)
    for k, p in enumerate(probs):
        c1 = solver.Context(p.tree, p.conn, p.M)
        v, bb, rr, vv = (c1.field(F64) for _ in range(4))
        bb.upload(p.b)
        solver.fill_ghost_all(v, p.bc)
        solver.vcycle(v, bb, rr, vv, 1.0e-3, None, None, geom, p.bc, solver.LaplacianDiscretization(2, 2),
                      ndown=1, ntop=2, nup=1)
        ref = p.b.like()
        v.download(ref)
        np.testing.assert_array_equal(_valid(got[k], p), _valid(ref.u, p))


@pytest.mark.parametrize("cfg", [4, 5, 6])
def test_forest_sibling_group_sweeps_equal_block_sweeps(cfg):
    probs = _problems(n_random=3)
    geom = solver.geometry_from_code(probs[0].geom)
    outs = []
    for c in (0, cfg):
        ctx, (u, b, r, u1) = _forest(probs)
        nl = max(len(p.tree) for p in probs)
        for l in range(1, nl + 1):
            ctx.set_sweep_config(F64, l, c)
        if c:
            print(f"cfg {cfg}: grouped levels {[ctx.level_grouped(l) for l in range(1, nl + 1)]}", flush=True)
        solver.fill_ghost_all(u, probs[0].bc)
        for _ in range(2):
            solver.vcycle_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc, [_disc(p) for p in probs],
                                  ndown=2, ntop=4, nup=2, coarse=solver.COARSE_EXACT)
        outs.append(np.asarray(_download(u, probs)[0]).copy())
    np.testing.assert_array_equal(outs[1], outs[0])


def _problems_3d():
    """Three 3D 8^3 trees with sibling groups on every refined level: a Poisson
    problem and two Helmholtz terms (per-tree block parameters)."""
    def crit(l, c):
        if l == 1:
            return (c[:, 0] >= 1) & (c[:, 1] <= 2) & (c[:, 2] <= 2)
        return (c[:, 0] <= 2) & (c[:, 1] <= 3)
    out = []
    for k, (helm, k2) in enumerate([(False, 0.0), (True, 3e4), (True, 4e5)]):
        out.append(configs.uniform_problem(3, (2, 2, 2), 8, 3, 0, ((1, 1), (1, 1), (-1, -1)), helm=helm, k2=k2,
                                           seed=20 + k, refine=crit))
    return out


@pytest.mark.parametrize("shape,order", [(7, 0), (8, 0), (8, 1)])
@pytest.mark.parametrize("mixed", [False, True])
def test_forest_3d_column_kernels_equal_single_solves(mixed, shape, order):
    # the column-mapped sweeps (shapes 7 and 8; the latter also visiting the
    # sibling groups in slab order, tree by tree), residual/restriction,
    # defect and norm kernels on forests: every tree gets its own bits
    probs = _problems_3d()
    geom = solver.CartesianGeometry()
    ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], 8)
    ctx.set_group_order(order)
    nl = max(len(p.tree) for p in probs)
    grouped = [ctx.level_grouped(l) for l in range(1, nl + 1)]
    print(f"grouped levels {grouped}", flush=True)
    assert all(grouped[1:]), "the column kernels run on grouped levels only"
    for l in range(1, nl + 1):
        ctx.set_sweep_config(F64, l, shape)
        ctx.set_sweep_config(F32, l, shape)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    b.upload(solver.concat_fields([p.b for p in probs]))
    discs = [_disc(p) for p in probs]
    if mixed:
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        n, hist = solver.solve_mixed_batched(u, b, e, be, re, ue, [p.s for p in probs], geom, probs[0].bc, discs,
                                             tol=1e-30, max_cycles=3, inner_cycles=2)
    else:
        n, hist = solver.solve_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc, discs, tol=1e-30,
                                       max_cycles=3)
    # a fixed number of cycles for every tree (the batch would otherwise run
    # until its slowest tree converges)
    assert n == 3
    got = _download(u, probs)
    for k, p in enumerate(probs):
        c1 = solver.Context(p.tree, p.conn, p.M)
        # the forest's three-launch level-1 solve of its stacked grids against
        # the axis-by-axis transforms of one grid
        c1.set_axis_gemm(1)
        for l in range(1, len(p.tree) + 1):
            c1.set_sweep_config(F64, l, shape)
            c1.set_sweep_config(F32, l, shape)
        v, bb, rr, vv = (c1.field(F64) for _ in range(4))
        bb.upload(p.b)
        if mixed:
            e1, be1, re1, ue1 = (c1.field(F32) for _ in range(4))
            n1, h1 = solver.solve_mixed(v, bb, e1, be1, re1, ue1, p.s, geom, p.bc, _disc(p), tol=1e-30,
                                        max_cycles=3, inner_cycles=2)
        else:
            n1, h1 = solver.solve(v, bb, rr, vv, p.s, geom, p.bc, _disc(p), tol=1e-30, max_cycles=3)
        assert n1 == 3
        # per-tree norms are summed in another fixed order than a lone tree's
        np.testing.assert_allclose(hist[k], h1, rtol=1e-12, atol=0)
        ref = p.b.like()
        v.download(ref)
        np.testing.assert_array_equal(_valid(got[k], p), _valid(ref.u, p))


def test_forest_autotune_covers_every_level_and_keeps_results():
    # the forest's first tree is the shallowest: autotune reports one shape per
    # level of the deepest tree, and tuned shapes give the same bits
    probs = sorted(_problems(n_random=3), key=lambda p: len(p.tree))
    assert len(probs[0].tree) < max(len(p.tree) for p in probs)
    geom = solver.geometry_from_code(probs[0].geom)
    discs = [_disc(p) for p in probs]
    outs = []
    for tune in (False, True):
        ctx, (u, b, r, u1) = _forest(probs)
        if tune:
            shapes = ctx.autotune(u, b, probs[0].s, geom, probs[0].bc, discs[0], reps=2)
            print(f"tuned shapes {shapes.tolist()}", flush=True)
            assert len(shapes) == ctx.nlevels == max(len(p.tree) for p in probs)
            b.upload(solver.concat_fields([p.b for p in probs]))
        n, hist = solver.solve_batched(u, b, r, u1, [p.s for p in probs], geom, probs[0].bc, discs, tol=1e-30,
                                       max_cycles=3)
        outs.append((hist.copy(), [_valid(g, p).copy() for g, p in zip(_download(u, probs), probs)]))
    np.testing.assert_array_equal(outs[1][0], outs[0][0])
    for a, c in zip(outs[1][1], outs[0][1]):
        np.testing.assert_array_equal(a, c)
