"""Order-4 V-cycles on the device (LaplacianDiscretization{D,4}, Cartesian,
complete levels: the four-colour smoother, the 9- and 19-point operator with
its corner and edge ghosts) against the oracle, bit for bit; the right-hand
side R q first; trees with coarse-fine faces are refused."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

from test_order4 import complete_problem, order4_oracle
from srmg import configs, solver
from srmg.solver import F64

pytestmark = pytest.mark.gpu

CASES = {
    "2d-m16": lambda: complete_problem(2, (2, 3), 16, levels=3, bcd=((1, -1), (-1, 1))),
    "2d-m8": lambda: complete_problem(2, (3, 2), 8, levels=2, bcd=((-1, -1), (1, 1))),
    "3d-m8": lambda: complete_problem(3, (2, 2, 3), 8, levels=2, bcd=((1, 1), (-1, -1), (-1, 1))),
}
GEOM = {"2d-m16": 0, "2d-m8": 0, "3d-m8": 0, "cyl1-m16": 1, "cyl2-m8": 2}
CASES["cyl1-m16"] = lambda: complete_problem(2, (2, 3), 16, levels=2, bcd=((1, -1), (-1, -1)))
CASES["cyl2-m8"] = lambda: complete_problem(2, (3, 2), 8, levels=3, bcd=((-1, -1), (1, -1)))


def _oracle(p, name):
    from oracle import Oracle
    from srmg.tree import box_stencil, connectivity
    return Oracle(p.tree, connectivity(p.tree, box_stencil(p.D)), p.M, p.b.G, GEOM[name], p.bc, order=4)


@pytest.mark.parametrize("name", list(CASES))
def test_order4_vcycles_match_oracle(name):
    p = CASES[name]()
    print(f"case {name}: {p.tree.nblocks()} blocks", flush=True)
    geom, disc = solver.geometry_from_code(GEOM[name]), solver.LaplacianDiscretization(p.D, 4)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
    q.upload(p.b)
    solver.rhs(q1, q, geom, p.bc, disc)
    for _ in range(3):
        solver.fill_ghost_all(u, p.bc, disc)
        solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    got, gb = p.b.like(), p.b.like()
    u.download(got)
    q1.download(gb)
    o = _oracle(p, name)
    uo, qo, q1o, ro, u1o = p.b.like(), p.b.copy(), p.b.like(), p.b.like(), p.b.like()
    o.rhs(q1o, qo, order=4)
    for _ in range(3):
        o.fill_ghost_all(uo)
        o.vcycle(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * p.D
    np.testing.assert_array_equal(got.u[vs], uo.u[vs])
    np.testing.assert_array_equal(gb.u[vs], q1o.u[vs])  # the FAS right-hand side too
    assert np.max(np.abs(uo.u[vs])) > 0


AMR = {
    "amr2-cyl1": lambda: (__import__("regrid_cases").problem(2), 1),
    "amr3": lambda: (__import__("regrid_cases").problem(3), 0),
    "cfg1-cart": lambda: (configs.make_cfg1(0), 0),
}


@pytest.mark.parametrize("name", list(AMR))
def test_order4_vcycles_on_amr_trees_match_oracle(name):
    # coarse-fine faces, edges and corners: the explicit ghost frames filled in
    # the reference's order (copy, boundary, interpolation), warm restarts
    p, gcode = AMR[name]()
    print(f"case {name}: {p.tree.nblocks()} blocks, levels {[len(l.pairs) for l in p.tree]}", flush=True)
    geom, disc = solver.geometry_from_code(gcode), solver.LaplacianDiscretization(p.D, 4)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
    q.upload(p.b)
    solver.rhs(q1, q, geom, p.bc, disc)
    for _ in range(3):
        solver.fill_ghost_all(u, p.bc, disc)
        solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    got, gb = p.b.like(), p.b.like()
    u.download(got)
    q1.download(gb)
    from oracle import Oracle
    from srmg.tree import box_stencil, connectivity
    o = Oracle(p.tree, connectivity(p.tree, box_stencil(p.D)), p.M, p.b.G, gcode, p.bc, order=4)
    uo, qo, q1o, ro, u1o = p.b.like(), p.b.copy(), p.b.like(), p.b.like(), p.b.like()
    o.rhs(q1o, qo, order=4)
    for _ in range(3):
        o.fill_ghost_all(uo)
        o.vcycle(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * p.D
    np.testing.assert_array_equal(got.u[vs], uo.u[vs])
    np.testing.assert_array_equal(gb.u[vs], q1o.u[vs])


def test_order4_refuses_trees_not_nested_for_the_box_stencil():
    def crit(l, c):
        return (c[:, 0] >= 2) & (c[:, 0] <= 4) & (c[:, 1] >= 2) & (c[:, 1] <= 5 * l)
    p = configs.uniform_problem(2, (4, 4), 16, 3, 0, ((1, -1), (-1, -1)), refine=crit)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    with pytest.raises(solver.SrmgError, match="nesting"):
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, solver.CartesianGeometry(), p.bc,
                      solver.LaplacianDiscretization(2, 4))


@pytest.mark.parametrize("name", ["2d-m16", "3d-m8"])
def test_order4_resnorm_matches_oracle(name):
    # resnorm! with the order-4 operator after each V-cycle, against the oracle
    # to 1e-12 (the leaf sum runs in another fixed order); with the
    # reference's few level-1 sweeps the norm falls slowly on these trees
    p = CASES[name]()
    geom, disc = solver.CartesianGeometry(), solver.LaplacianDiscretization(p.D, 4)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
    q.upload(p.b)
    solver.rhs(q1, q, geom, p.bc, disc)
    o = order4_oracle(p)
    uo, qo, q1o, ro, u1o = p.b.like(), p.b.copy(), p.b.like(), p.b.like(), p.b.like()
    o.rhs(q1o, qo, order=4)
    hist = []
    for k in range(4):
        if k:
            solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
            o.fill_ghost_all(uo)
            o.vcycle(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
        got = solver.resnorm(r, q1, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc)
        o.fill_ghost_all(uo)
        want = o.resnorm(ro, q1o, uo, p.s)
        print(f"{name} cycle {k}: {got:.6e} (oracle {want:.6e})", flush=True)
        assert abs(got - want) <= 1e-12 * want
        hist.append(got)
    assert hist[-1] < hist[0]


def test_order4_poisson_step_matches_oracle():
    # derivs!'s Poisson block with the order-4 discretization on an AMR
    # cylindrical tree: restrict_full!, R, then warm-started cycles; two steps
    p, gcode = AMR["amr2-cyl1"]()
    geom, disc = solver.geometry_from_code(gcode), solver.LaplacianDiscretization(2, 4)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
    for _ in range(2):
        q.upload(p.b)
        solver.poisson_step(u, q1, r, u1, p.s, geom, p.bc, disc, poisson_iter=2, q=q)
    got = p.b.like()
    u.download(got)
    from oracle import Oracle
    from srmg.tree import box_stencil, connectivity
    o = Oracle(p.tree, connectivity(p.tree, box_stencil(2)), p.M, p.b.G, gcode, p.bc, order=4)
    uo, ro, u1o = p.b.like(), p.b.like(), p.b.like()
    for _ in range(2):
        qo, q1o = p.b.copy(), p.b.like()
        o.restrict_full(qo)
        o.rhs(q1o, qo, order=4)
        for _ in range(2):
            o.fill_ghost_all(uo)
            o.vcycle(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * 2
    np.testing.assert_array_equal(got.u[vs], uo.u[vs])


@pytest.mark.parametrize("name", ["amr2-cyl1", "amr3"])
def test_order4_fmg_matches_oracle(name):
    # fmg! with the order-4 operator (omega = 1.25, the reference's default),
    # twice from the previous result
    p, gcode = AMR[name]()
    geom, disc = solver.geometry_from_code(gcode), solver.LaplacianDiscretization(p.D, 4)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
    q.upload(p.b)
    solver.rhs(q1, q, geom, p.bc, disc)
    for _ in range(2):
        solver.fill_ghost_all(u, p.bc, disc)
        solver.fmg(u, q1, r, u1, p.s, geometry=geom, bc=p.bc, lpldisc=disc, ndown=2, ntop=4, nup=2)
    got = p.b.like()
    u.download(got)
    from oracle import Oracle
    from srmg.tree import box_stencil, connectivity
    o = Oracle(p.tree, connectivity(p.tree, box_stencil(p.D)), p.M, p.b.G, gcode, p.bc, order=4)
    uo, qo, q1o, ro, u1o = p.b.like(), p.b.copy(), p.b.like(), p.b.like(), p.b.like()
    o.rhs(q1o, qo, order=4)
    for _ in range(2):
        o.fill_ghost_all(uo)
        o.fmg(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * p.D
    np.testing.assert_array_equal(got.u[vs], uo.u[vs])


@pytest.mark.parametrize("name", ["amr2-cyl1", "amr3"])
def test_order4_solve_equals_its_cycles(name):
    # srmg_solve with the order-4 operator: the pre-fill, then captured cycles
    # and resnorm!; the same bits as the same calls one by one, and its history
    p, gcode = AMR[name]()
    geom, disc = solver.geometry_from_code(gcode), solver.LaplacianDiscretization(p.D, 4)
    out = []
    for mode in ("solve", "calls"):
        ctx = solver.Context(p.tree, p.conn, p.M)
        u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
        q.upload(p.b)
        solver.rhs(q1, q, geom, p.bc, disc)
        if mode == "solve":
            n, hist = solver.solve(u, q1, r, u1, p.s, geom, p.bc, disc, tol=1e-30, max_cycles=3,
                                   coarse=solver.COARSE_GS, ndown=2, ntop=4, nup=2)
            assert n == 3
        else:
            zero = ctx.field(F64)
            solver.fill_ghost_all(zero, p.bc, disc)
            r0 = solver.resnorm(r, q1, zero, p.s, geometry=geom, bc=p.bc, lpldisc=disc)
            solver.fill_ghost_all(u, p.bc, disc)
            hist = [solver.resnorm(r, q1, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc) / r0]
            for _ in range(3):
                solver.vcycle(u, q1, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
                hist.append(solver.resnorm(r, q1, u, p.s, geometry=geom, bc=p.bc, lpldisc=disc) / r0)
        host = p.b.like()
        u.download(host)
        out.append((host.u.copy(), np.asarray(hist)))
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * p.D
    np.testing.assert_array_equal(out[0][0][vs], out[1][0][vs])
    np.testing.assert_array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("name", ["amr2-cyl1", "amr3"])
def test_order4_mixed_precision_reaches_the_fp64_solution(name):
    # FP64 defect correction around FP32 order-4 V-cycles: to the same
    # tolerance as the FP64 order-4 solve, the potentials agree to 1e-6
    p, gcode = AMR[name]()
    geom, disc = solver.geometry_from_code(gcode), solver.LaplacianDiscretization(p.D, 4)
    sols = []
    for mixed in (False, True):
        ctx = solver.Context(p.tree, p.conn, p.M)
        u, q, q1, r, u1 = (ctx.field(F64) for _ in range(5))
        q.upload(p.b)
        solver.rhs(q1, q, geom, p.bc, disc)
        # the reference's level-1 sweeps converge slowly on the 48 x 64 cylindrical level 1
        tol, ntop = (1e-7, 100) if p.D == 2 else (1e-9, 30)
        kw = dict(tol=tol, max_cycles=400, coarse=solver.COARSE_GS, ndown=2, ntop=ntop, nup=2)
        if mixed:
            e, be, re, ue = (ctx.field(solver.F32) for _ in range(4))
            n, hist = solver.solve_mixed(u, q1, e, be, re, ue, p.s, geom, p.bc, disc, inner_cycles=2, **kw)
        else:
            n, hist = solver.solve(u, q1, r, u1, p.s, geom, p.bc, disc, **kw)
        print(f"{name} mixed={mixed}: {n} cycles, final {hist[-1]:.3e}", flush=True)
        assert hist[-1] <= tol
        host = p.b.like()
        u.download(host)
        G, M = p.b.G, p.M
        sols.append(host.u[(slice(G, G + M),) * p.D].copy())
    err = np.max(np.abs(sols[1] - sols[0])) / np.max(np.abs(sols[0]))
    print(f"{name}: potential rel diff {err:.2e}")
    assert err <= (1e-5 if p.D == 2 else 1e-6)
