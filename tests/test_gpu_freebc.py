"""Free boundary conditions on the device (srmg_freebc_*) against the oracle:
the boundary values, the sources added to q, the warm-start sequence of
setfreebc! and the electric field with the free-boundary ghosts, bit for bit
(the device sums the same products in the same order)."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

from oracle import Oracle
from srmg import configs, solver
from srmg.freebc import FreeBC
from srmg.solver import F64

pytestmark = pytest.mark.gpu


def _outer_refined():
    def crit(l, c):
        if l == 1:
            return (c[:, 0] >= 2) & (c[:, 1] >= 2) & (c[:, 1] <= 3)
        return (c[:, 0] == 6) & (c[:, 1] == 5)
    p = configs.uniform_problem(2, (3, 4), 16, 3, 1, ((1, -1), (-1, -1)), refine=crit, seed=4)
    return p, (3, 4)


def _cfg1():
    return configs.make_cfg1(2), (4, 8)


CASES = {"outer-refined": _outer_refined, "cfg1": _cfg1}


def _valid(f, p):
    G, M = f.G, p.M
    return f.u[(slice(G, G + M),) * p.D]


@pytest.mark.parametrize("name", list(CASES))
def test_freebc_apply_matches_oracle(name):
    p, root = CASES[name]()
    level = len(p.tree)
    print(f"case {name}: levels {level}, blocks {p.tree.nblocks()}", flush=True)
    ctx = solver.Context(p.tree, p.conn, p.M)
    fb = FreeBC(ctx, solver.CylindricalGeometry(1), level, root, p.h)
    u_host = p.b.copy()  # any potential: the outer cells are what matters
    q_host = p.b.copy()
    q_host.u *= 0.5
    u, q = ctx.field(F64), ctx.field(F64)
    u.upload(u_host)
    q.upload(q_host)
    fb.freebc(q, u, p.s)
    b = fb.boundary()
    got = p.b.like()
    q.download(got)
    o = Oracle.for_problem(p)
    qo = q_host.copy()
    bo = o.freebc(qo, u_host, p.h, p.s, 1, level, root, fb.f)
    np.testing.assert_array_equal(b, bo)
    np.testing.assert_array_equal(_valid(got, p), _valid(qo, p))
    assert not np.array_equal(_valid(qo, p), _valid(q_host, p))


@pytest.mark.parametrize("name", list(CASES))
def test_setfreebc_sequence_matches_oracle(name):
    # setfreebc!: poisson_iter warm-start V-cycles on the homogeneous potential
    # (the reference's level-1 sweeps), then freebc! into q1
    p, root = CASES[name]()
    level = len(p.tree)
    geom, disc = solver.CylindricalGeometry(1), solver.LaplacianDiscretization(2, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    fb = FreeBC(ctx, geom, level, root, p.h)
    q1, r, u1 = (ctx.field(F64) for _ in range(3))
    q1.upload(p.b)
    fb.setfreebc(q1, r, u1, p.s, p.bc, disc, poisson_iter=2, nup=2, ndown=2, ntop=4)
    gq, gu = p.b.like(), p.b.like()
    q1.download(gq)
    fb.uh.download(gu)
    o = Oracle.for_problem(p)
    uh, qo, ro, u1o = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    for _ in range(2):
        o.fill_ghost_all(uh)
        o.vcycle(uh, qo, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    o.freebc(qo, uh, p.h, p.s, 1, level, root, fb.f)
    np.testing.assert_array_equal(_valid(gu, p), _valid(uh, p))
    np.testing.assert_array_equal(_valid(gq, p), _valid(qo, p))


@pytest.mark.parametrize("name", list(CASES))
def test_electric_field_with_free_ghosts_matches_oracle(name):
    # the post-solve ghost fill with fill_ghost_free! (src/streamer.jl:340-349)
    p, root = CASES[name]()
    level = len(p.tree)
    geom, disc = solver.CylindricalGeometry(1), solver.LaplacianDiscretization(2, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    fb = FreeBC(ctx, geom, level, root, p.h)
    q1, r, u1, u = (ctx.field(F64) for _ in range(4))
    q1.upload(p.b)
    fb.setfreebc(q1, r, u1, p.s, p.bc, disc, poisson_iter=1)
    solver.fill_ghost_all(u, p.bc)
    for _ in range(2):
        solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    e0 = [0.0, -2.0e6]
    eabs = ctx.field(F64)
    E = fb.electric_field(u, eabs, p.h, e0, p.bc, G=p.b.G)
    b = fb.boundary()
    host = p.b.like()
    u.download(host)
    o = Oracle.for_problem(p)
    o.fill_ghost_all(host)
    for l in range(1, level + 1):
        o.fill_ghost_free(host, b, 1, level, l, root)
    Eo, fabs = o.electric_field(host, p.h, e0)
    np.testing.assert_array_equal(E, Eo)
    got = p.b.like()
    eabs.download(got)
    np.testing.assert_array_equal(_valid(got, p), fabs[(slice(p.b.G, p.b.G + p.M),) * p.D])
    # the free ghosts changed the radial field at the outer boundary
    Ep = solver.electric_field(u, eabs, p.h, e0, p.bc, G=p.b.G)
    assert not np.array_equal(E, Ep)


@pytest.mark.parametrize("poisson_fmg", [False, True])
def test_poisson_step_with_free_boundary_matches_oracle(poisson_fmg):
    # derivs!'s Poisson block (src/streamer.jl:320-333): setfreebc!, then
    # poisson_iter warm-start cycles (fmg! or vcycle!) on u
    p, root = _outer_refined()
    level = len(p.tree)
    geom, disc = solver.CylindricalGeometry(1), solver.LaplacianDiscretization(2, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    fb = FreeBC(ctx, geom, level, root, p.h)
    u, q1, r, u1 = (ctx.field(F64) for _ in range(4))
    for _ in range(2):  # two steps: the second starts warm (u and the homogeneous potential)
        q1.upload(p.b)  # the charge, as rhs_level! rebuilds q1 every step
        solver.poisson_step(u, q1, r, u1, p.s, geom, p.bc, disc, poisson_iter=2, poisson_fmg=poisson_fmg,
                            freebc=fb)
    got = p.b.like()
    u.download(got)
    o = Oracle.for_problem(p)
    uo, uh, ro, u1o = p.b.like(), p.b.like(), p.b.like(), p.b.like()
    for _ in range(2):
        qo = p.b.copy()
        for _ in range(2):
            o.fill_ghost_all(uh)
            if poisson_fmg:
                o.fmg(uh, qo, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
            else:
                o.vcycle(uh, qo, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
        o.freebc(qo, uh, p.h, p.s, 1, level, root, fb.f)
        for _ in range(2):
            o.fill_ghost_all(uo)
            if poisson_fmg:
                o.fmg(uo, qo, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
            else:
                o.vcycle(uo, qo, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    np.testing.assert_array_equal(_valid(got, p), _valid(uo, p))


def test_freebc_refuses_a_mismatched_setup():
    p, root = _outer_refined()
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q = ctx.field(F64), ctx.field(F64)
    u.upload(p.b)
    q.upload(p.b)
    geom = solver.CylindricalGeometry(1)
    # the tree is finer than the level of the boundary (src/freebc.jl:94-95)
    with pytest.raises(solver.SrmgError, match="finest level"):
        FreeBC(ctx, geom, len(p.tree) - 1, root, p.h).freebc(q, u, p.s)
    with pytest.raises(solver.SrmgError, match="root domain"):
        FreeBC(ctx, geom, len(p.tree), (root[0] + 1, root[1]), p.h).freebc(q, u, p.s)
    with pytest.raises(solver.SrmgError, match="2D cylindrical"):
        FreeBC(ctx, solver.CartesianGeometry(), len(p.tree), root, p.h)
