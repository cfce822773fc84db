"""GPU parity: the HIP library through its C ABI against the CPU oracle.

The FP64 device path performs the same floating-point operations in the same
order as the oracle, so potentials and FAS right-hand sides must agree bit for
bit, with the reference's coarsest-level treatment (ntop Gauss-Seidel sweeps)
and with the exact coarsest-level solve alike (both sides diagonalise the
level-1 operator with the same implicit QL method and apply the transforms as
k-ordered sums). Residual norms are summed in another fixed order on the
device: there the bar is 1e-12 relative (histories: 1e-12 of resnorm(0)).
"""

import numpy as np
import pytest

# torch's HIP runtime must be the process's first: import it before the
# library is loaded (a later first import of torch finds no GPU)
torch = pytest.importorskip("torch")

from oracle import Oracle
from srmg import configs, solver
from srmg.solver import F32, F64

pytestmark = pytest.mark.gpu


def _tiny(D, geom, bcd, helm=False, k2=0.0, M=None, seed=0):
    if D == 2:
        rootsize = (3, 4)
        M = M or 16

        def crit(l, c):
            if l == 1:
                return (c[:, 0] <= 2) & (c[:, 1] >= 2) & (c[:, 1] <= 3)
            return (c[:, 0] == 2) & (c[:, 1] == 5)
        levels = 3
    else:
        rootsize = (2, 3, 2)
        M = M or 8

        def crit(l, c):
            if l == 1:
                return (c[:, 0] == 1) & (c[:, 1] == 2)
            return (c[:, 0] == 2) & (c[:, 1] == 3) & (c[:, 2] == 2)
        levels = 3
    return configs.uniform_problem(D, rootsize, M, levels, geom, bcd, helm=helm, k2=k2, seed=seed,
                                   refine=crit)


CASES = {
    "cfg1": lambda: configs.make_cfg1(3),
    "cyl1-m16": lambda: _tiny(2, 1, ((1, -1), (-1, -1))),
    "cyl2-m8": lambda: _tiny(2, 2, ((-1, 1), (1, -1)), M=8),
    "cart2-helm": lambda: _tiny(2, 0, ((1, 1), (-1, -1)), helm=True, k2=4e5),
    "cart3-m8": lambda: _tiny(3, 0, ((1, 1), (1, 1), (-1, -1))),
    "cart3-helm": lambda: _tiny(3, 0, ((-1, -1), (1, 1), (-1, 1)), helm=True, k2=2e5),
    "cart3-m16": lambda: _tiny(3, 0, ((1, -1), (1, 1), (-1, -1)), M=16),
    "cfg3-term2": lambda: configs.make_helmholtz_cfg1(0)[2],
}


def _device(p, dtype=F64):
    ctx = solver.Context(p.tree, p.conn, p.M)
    return ctx, [ctx.field(dtype) for _ in range(4)]


def _valid(f, p):
    G, M = f.G, p.M
    return f.u[(slice(G, G + M),) * p.D]


def _disc(p):
    return (solver.HelmholtzDiscretization(p.D, 2, p.k2) if p.helm
            else solver.LaplacianDiscretization(p.D, 2))


@pytest.mark.parametrize("name", list(CASES))
def test_upload_download_roundtrip(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    got = p.b.like()
    b.download(got)
    np.testing.assert_array_equal(_valid(got, p), _valid(p.b, p))


@pytest.mark.parametrize("name", list(CASES))
def test_fill_ghost_matches_oracle(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    rng = np.random.default_rng(7)
    host = p.b.like()
    host.u[...] = rng.standard_normal(host.u.shape)
    o = Oracle.for_problem(p)
    o.fill_ghost_all(host)
    ctx, (u, b, r, u1) = _device(p)
    u.upload(host)
    solver.fill_ghost_all(u, p.bc)
    got = p.b.like()
    got.u[...] = np.nan
    u.download(got, bc=p.bc)
    G, M, D = p.b.G, p.M, p.D
    # depth-1 face ghosts of every block
    for layer in p.tree:
        for c, blk in layer.pairs:
            for d in range(D):
                for side, gi in ((0, G - 1), (1, G + M)):
                    sl = [slice(G, G + M)] * D
                    sl[d] = gi
                    sl = tuple(sl) + (blk - 1,)
                    np.testing.assert_array_equal(got.u[sl], host.u[sl])


def _run_cycles(p, n, fmg=False, coarse=solver.COARSE_GS, opts=(2, 4, 2), omega=1.0):
    ndown, ntop, nup = opts
    o = Oracle.for_problem(p)
    ou, ob, orr, ou1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    o.fill_ghost_all(ou)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    solver.fill_ghost_all(u, p.bc)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    exact = coarse == solver.COARSE_EXACT
    for _ in range(n):
        if fmg:
            o.fmg(ou, ob, orr, ou1, p.s, ndown=ndown, ntop=ntop, nup=nup, omega=omega, coarse_exact=exact)
            solver.fmg(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=ndown, ntop=ntop,
                       nup=nup, omega=omega, coarse=coarse)
        else:
            o.vcycle(ou, ob, orr, ou1, p.s, ndown=ndown, ntop=ntop, nup=nup, omega=omega,
                     coarse_exact=exact)
            solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=ndown, ntop=ntop,
                          nup=nup, omega=omega, coarse=coarse)
    gu, gb = p.b.like(), p.b.like()
    u.download(gu)
    b.download(gb)
    return (gu, gb), (ou, ob), (ctx, u, b, r, u1, o, orr)


@pytest.mark.parametrize("name", list(CASES))
def test_vcycle_bitwise_reference_mode(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    (gu, gb), (ou, ob), _ = _run_cycles(p, 4)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))
    # b is mutated at covered coarse cells with the FAS right-hand side
    np.testing.assert_array_equal(_valid(gb, p), _valid(ob, p))


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8", "cart2-helm"])
def test_vcycle_bitwise_other_options(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    (gu, gb), (ou, ob), _ = _run_cycles(p, 2, opts=(3, 5, 1), omega=1.15)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8", "cart3-helm"])
def test_fmg_bitwise(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    (gu, gb), (ou, ob), _ = _run_cycles(p, 1, fmg=True, omega=1.25)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))
    np.testing.assert_array_equal(_valid(gb, p), _valid(ob, p))


@pytest.mark.parametrize("name", list(CASES))
def test_resnorm_matches_oracle(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    (gu, gb), (ou, ob), (ctx, u, b, r, u1, o, orr) = _run_cycles(p, 2)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    g = solver.resnorm(r, b, u, p.s, p.tree, p.conn, geom, p.bc, disc)
    e = o.resnorm(orr, ob, ou, p.s)
    assert g == pytest.approx(e, rel=1e-12)


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8", "cart3-helm", "cfg3-term2"])
def test_exact_coarse_solve_history(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    o = Oracle.for_problem(p)
    ou, ob, orr, ou1 = p.b.like(), p.b.copy(), p.b.like(), p.b.like()
    r0 = o.resnorm(orr, ob, ou, p.s)
    o.fill_ghost_all(ou)
    ohist = [o.resnorm(orr, ob, ou, p.s) / r0]
    for _ in range(12):
        o.vcycle(ou, ob, orr, ou1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
        ohist.append(o.resnorm(orr, ob, ou, p.s) / r0)
        if ohist[-1] <= 1e-10:
            break
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    n, hist = solver.solve(u, b, r, u1, p.s, solver.geometry_from_code(p.geom), p.bc, _disc(p),
                           tol=1e-10, max_cycles=12)
    assert n == len(ohist) - 1
    np.testing.assert_allclose(hist, ohist, rtol=0, atol=1e-12)
    gu = p.b.like()
    u.download(gu)
    ref = _valid(ou, p)
    err = np.max(np.abs(_valid(gu, p) - ref)) / np.max(np.abs(ref))
    print(f"{name}: {n} cycles, potential max rel diff vs oracle {err:.3e}", flush=True)
    np.testing.assert_array_equal(_valid(gu, p), ref)


@pytest.mark.parametrize("name", ["cfg1", "cyl1-m16", "cyl2-m8", "cart2-helm", "cart3-m8", "cart3-helm"])
def test_vcycle_and_fmg_bitwise_exact_coarse(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    (gu, gb), (ou, ob), _ = _run_cycles(p, 3, coarse=solver.COARSE_EXACT)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))
    np.testing.assert_array_equal(_valid(gb, p), _valid(ob, p))
    (gu, gb), (ou, ob), _ = _run_cycles(p, 1, fmg=True, omega=1.25, coarse=solver.COARSE_EXACT)
    np.testing.assert_array_equal(_valid(gu, p), _valid(ou, p))
    np.testing.assert_array_equal(_valid(gb, p), _valid(ob, p))


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8", "cfg4-small"])
def test_axis_transform_variants_give_identical_solves(name):
    """The three-launch level-1 solve (variant 2, the default), the one-wave
    transform kernels (1) and the LDS-tiled ones (0) sum the same products in
    the same order: exact-coarse solves agree bit for bit."""
    p = configs.make_cfg4(0, scale=0.35, levels=3) if name == "cfg4-small" else CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    res = []
    for variant in (0, 1, 2):
        ctx, (u, b, r, u1) = _device(p)
        ctx.set_axis_gemm(variant)
        b.upload(p.b)
        n, hist = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-10, max_cycles=6)
        host = p.b.like()
        u.download(host)
        res.append((np.asarray(hist), _valid(host, p).copy()))
    for k in (1, 2):
        np.testing.assert_array_equal(res[k][0], res[0][0])
        np.testing.assert_array_equal(res[k][1], res[0][1])


@pytest.mark.parametrize("D,n", [(3, (64, 64, 64)), (3, (64, 32, 16)), (3, (16, 40, 64)), (3, (32, 16, 128)),
                                 (2, (64, 128)), (2, (128, 256)), (2, (96, 200))])
def test_axis_transform_variants_give_identical_grid_solves(D, n):
    p = CASES["cart3-m8" if D == 3 else "cyl2-m8"]()
    bc = p.bc
    geom = solver.CartesianGeometry()
    disc = solver.LaplacianDiscretization(D, 2)
    V = int(np.prod(n))
    g = torch.Generator(device="cpu").manual_seed(7)
    f = torch.randn((2, V), dtype=torch.float64, generator=g).cuda()
    outs = []
    for variant in (0, 1, 2):
        ctx = solver.Context(p.tree, p.conn, p.M)
        ctx.set_axis_gemm(variant)
        u = torch.empty_like(f)
        torch.cuda.synchronize()
        solver.grid_solve(ctx, D, n, 0.7, geom, bc, disc, f, u, 2)
        torch.cuda.synchronize()
        outs.append(u.cpu().numpy())
    np.testing.assert_array_equal(outs[1], outs[0])
    np.testing.assert_array_equal(outs[2], outs[0])


def test_cached_solve_graphs_replay_exactly():
    """A context keeps the captured step of each solve configuration; a later
    solve with the same configuration replays it from its first cycle, and a
    solve with another scale or shape captures its own."""
    p = configs.make_cfg4(0, scale=0.3, levels=3)
    geom, disc = solver.CartesianGeometry(), solver.LaplacianDiscretization(3, 2)

    def run(ctx, fs, s, mixed):
        u, b = fs[0], fs[1]
        b.upload(p.b)
        if mixed:
            n, h = solver.solve_mixed(u, b, *fs[2:], s, geom, p.bc, disc, tol=1e-9, max_cycles=12, inner_cycles=2,
                                      guess=solver.GUESS_ZERO)
        else:
            n, h = solver.solve(u, b, *fs[2:], s, geom, p.bc, disc, tol=1e-9, max_cycles=12,
                                guess=solver.GUESS_ZERO)
        host = p.b.like()
        u.download(host)
        return np.asarray(h).copy(), _valid(host, p).copy()

    for mixed in (False, True):
        ctx = solver.Context(p.tree, p.conn, p.M)
        fs = [ctx.field(F64), ctx.field(F64)] + [ctx.field(F32 if mixed else F64) for _ in range(4 if mixed else 2)]
        first = run(ctx, fs, p.s, mixed)
        again = run(ctx, fs, p.s, mixed)
        np.testing.assert_array_equal(again[0], first[0])
        np.testing.assert_array_equal(again[1], first[1])
        other = run(ctx, fs, 2.0 * p.s, mixed)
        ctx2 = solver.Context(p.tree, p.conn, p.M)
        fs2 = [ctx2.field(F64), ctx2.field(F64)] + [ctx2.field(F32 if mixed else F64)
                                                    for _ in range(4 if mixed else 2)]
        fresh = run(ctx2, fs2, 2.0 * p.s, mixed)
        np.testing.assert_array_equal(other[0], fresh[0])
        np.testing.assert_array_equal(other[1], fresh[1])
        assert not np.array_equal(other[1], first[1])
        for l in range(1, len(p.tree) + 1):
            ctx.set_sweep_config(F32 if mixed else F64, l, 0)
        shape0 = run(ctx, fs, p.s, mixed)
        np.testing.assert_array_equal(shape0[0], first[0])
        np.testing.assert_array_equal(shape0[1], first[1])


def test_cached_solve_graphs_after_regrid_and_operator_changes():
    """Captured steps never outlive what they point to: after a regrid, or a
    solve with other boundary conditions (which rebuilds the level-1 matrices),
    every solve equals the same solve in a fresh context."""
    pa = CASES["cart3-m8"]()
    pb = configs.make_cfg4(1, scale=0.3, levels=3)
    geom, disc = solver.CartesianGeometry(), solver.LaplacianDiscretization(3, 2)
    bc2 = configs.boundaryconditions(((-1, -1), (1, 1), (-1, 1)))

    def solve(ctx, p, bc):
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        b.upload(p.b)
        n, h = solver.solve(u, b, r, u1, p.s, geom, bc, disc, tol=1e-9, max_cycles=6, guess=solver.GUESS_ZERO)
        host = p.b.like()
        u.download(host)
        for f in (u, b, r, u1):
            f.close()
        return np.asarray(h).copy(), _valid(host, p).copy()

    def fresh(p, bc):
        return solve(solver.Context(p.tree, p.conn, p.M), p, bc)

    ctx = solver.Context(pa.tree, pa.conn, pa.M)
    for step, (p, bc) in enumerate([(pa, pa.bc), (pa, pa.bc), (pa, bc2), (pa, pa.bc), (pb, pb.bc), (pa, pa.bc)]):
        if ctx.tree is not p.tree:
            ctx.regrid(p.tree, p.conn)
        got = solve(ctx, p, bc)
        ref = fresh(p, bc)
        print(f"step {step}: cycles {len(got[0]) - 1}", flush=True)
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])


@pytest.mark.parametrize("use_graph", [False, True])
def test_solve_graph_replay_matches_eager(use_graph):
    p = configs.make_cfg1(1)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    n, hist = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-9, use_graph=use_graph)
    ctx2, (v, b2, r2, v1) = _device(p)
    b2.upload(p.b)
    for _ in range(n):
        solver.vcycle(v, b2, r2, v1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2,
                      coarse=solver.COARSE_EXACT)
    a, e = p.b.like(), p.b.like()
    u.download(a)
    v.download(e)
    np.testing.assert_array_equal(_valid(a, p), _valid(e, p))


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8", "cfg3-term2"])
def test_mixed_precision_matches_fp64(name):
    p = CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    n64, h64 = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-8)
    ctx2 = solver.Context(p.tree, p.conn, p.M)
    v, b2 = ctx2.field(F64), ctx2.field(F64)
    e, be, re, ue = (ctx2.field(F32) for _ in range(4))
    b2.upload(p.b)
    n32, h32 = solver.solve_mixed(v, b2, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-8)
    assert h32[-1] <= 1e-8 and h64[-1] <= 1e-8
    a, c = p.b.like(), p.b.like()
    u.download(a)
    v.download(c)
    ref = _valid(a, p)
    err = np.max(np.abs(_valid(c, p) - ref)) / np.max(np.abs(ref))
    print(f"{name}: fp64 {n64} cycles, mixed {n32} cycles, potential rel err {err:.2e}")
    assert err <= 1e-7


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8"])
def test_mixed_first_correction_cycle_count(name):
    """inner_first sets the FP32 V-cycles of the first correction only: the
    first outer step of (inner_first=a, inner_cycles=b) is that of
    inner_cycles=a, bit for bit, and inner_first=b is inner_cycles=b; the
    captured graphs of both steps replay to the same bits."""
    p = CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx = solver.Context(p.tree, p.conn, p.M)
    v, b = ctx.field(F64), ctx.field(F64)
    e, be, re, ue = (ctx.field(F32) for _ in range(4))
    b.upload(p.b)

    def run(inner, first=0):
        n, h = solver.solve_mixed(v, b, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-8, max_cycles=12,
                                  inner_cycles=inner, inner_first=first, guess=solver.GUESS_ZERO)
        host = p.b.like()
        v.download(host)
        return n, h, _valid(host, p).copy()

    n2, h2, _ = run(2)
    n3, h3, u3 = run(3)
    n23, h23, u23 = run(3, 2)
    n32, h32, _ = run(2, 3)
    n33, h33, u33 = run(3, 3)
    print(f"{name}: [2] {h2}, [3] {h3}, [2,3] {h23}, [3,2] {h32}", flush=True)
    assert h23[1] == h2[1] and h32[1] == h3[1]
    assert h23[-1] <= 1e-8 and h32[-1] <= 1e-8
    assert n33 == n3
    np.testing.assert_array_equal(h33, h3)
    np.testing.assert_array_equal(u33, u3)
    # second solve of the same schedule: both graphs replayed
    m23, g23, w23 = run(3, 2)
    assert m23 == n23
    np.testing.assert_array_equal(g23, h23)
    np.testing.assert_array_equal(w23, u23)
    err = np.max(np.abs(u23 - u3)) / np.max(np.abs(u3))
    assert err <= 1e-7, err


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8", "cart3-m16"])
@pytest.mark.parametrize("dtype", [F64, F32])
def test_every_sweep_tile_shape_gives_identical_results(name, dtype):
    p = CASES[name]()
    print(f"case {name} dtype {dtype}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    outs = []
    grouped = []
    for cfg in range(solver.SWEEP_SHAPES):
        ctx, (u, b, r, u1) = _device(p, dtype)
        if cfg == 0:
            grouped = [ctx.level_grouped(l) for l in range(1, len(p.tree) + 1)]
            print(f"  levels with sibling groups: {grouped}", flush=True)
        for l in range(1, len(p.tree) + 1):
            ctx.set_sweep_config(dtype, l, cfg)
        b.upload(p.b)
        solver.fill_ghost_all(u, p.bc)
        for _ in range(2):
            solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
        host = p.b.like()
        u.download(host)
        outs.append(_valid(host, p).copy())
    for o in outs[1:]:
        np.testing.assert_array_equal(o, outs[0])


@pytest.mark.parametrize("name", ["cart3-m8", "cfg4-small"])
def test_group_order_gives_identical_solves(name):
    """The column sweeps visit sibling groups in slab order or in block order:
    the same bits (FP64 and mixed solves, every sweep variant of shapes 8-11)."""
    p = configs.make_cfg4(0, scale=0.35, levels=3) if name == "cfg4-small" else CASES[name]()
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    for cfg in (8, 9, 10, 11):
        res = []
        for order in (0, 1):
            ctx = solver.Context(p.tree, p.conn, p.M)
            ctx.set_group_order(order)
            for l in range(1, len(p.tree) + 1):
                ctx.set_sweep_config(F64, l, cfg)
                ctx.set_sweep_config(F32, l, cfg)
            u, b, r, u1 = (ctx.field(F64) for _ in range(4))
            b.upload(p.b)
            n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12)
            v, b2 = ctx.field(F64), ctx.field(F64)
            b2.upload(p.b)
            e, be, re, ue = (ctx.field(F32) for _ in range(4))
            nm, hm = solver.solve_mixed(v, b2, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12,
                                        inner_cycles=3, inner_first=2)
            outs = []
            for f in (u, v):
                host = p.b.like()
                f.download(host)
                outs.append(_valid(host, p).copy())
            res.append((h, hm, outs))
        print(f"{name} shape {cfg}: fp64 {len(res[0][0])} cycles, mixed {len(res[0][1])} steps", flush=True)
        np.testing.assert_array_equal(res[1][0], res[0][0])
        np.testing.assert_array_equal(res[1][1], res[0][1])
        np.testing.assert_array_equal(res[1][2][0], res[0][2][0])
        np.testing.assert_array_equal(res[1][2][1], res[0][2][1])


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8", "cfg4-small"])
def test_sibling_group_kernels_give_identical_solves(name):
    p = configs.make_cfg4(0, scale=0.35, levels=3) if name == "cfg4-small" else CASES[name]()
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    res = []
    for on in (False, True):
        ctx = solver.Context(p.tree, p.conn, p.M)
        ctx.set_group_tiles(on)
        if on:
            print(f"{name}: grouped levels {[ctx.level_grouped(l) for l in range(1, len(p.tree) + 1)]}", flush=True)
            for l in range(1, len(p.tree) + 1):
                ctx.set_sweep_config(F64, l, 5)
                ctx.set_sweep_config(F32, l, 6)
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        b.upload(p.b)
        n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12)
        v, b2 = ctx.field(F64), ctx.field(F64)
        b2.upload(p.b)
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        nm, hm = solver.solve_mixed(v, b2, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12,
                                    inner_cycles=2)
        outs = []
        for f in (u, v):
            host = p.b.like()
            f.download(host)
            outs.append(_valid(host, p).copy())
        res.append((h, hm, outs))
    np.testing.assert_array_equal(res[1][0], res[0][0])
    np.testing.assert_array_equal(res[1][1], res[0][1])
    np.testing.assert_array_equal(res[1][2][0], res[0][2][0])
    np.testing.assert_array_equal(res[1][2][1], res[0][2][1])


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8", "cart3-m16"])
def test_electric_field_matches_oracle(name):
    # electric_field! after two V-cycles: E on the faces and |E| at the centres,
    # bit for bit against the oracle on the same potential
    p = CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    solver.fill_ghost_all(u, p.bc)
    for _ in range(2):
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    e0 = [0.0, -2.0e6, 1.0e5][:p.D]
    eabs = ctx.field(F64)
    E = solver.electric_field(u, eabs, p.h, e0, p.bc, G=p.b.G)
    host = p.b.like()
    u.download(host)
    o = Oracle.for_problem(p)
    o.fill_ghost_all(host)
    Eo, fabs = o.electric_field(host, p.h, e0)
    np.testing.assert_array_equal(E, Eo)
    got = p.b.like()
    eabs.download(got)
    np.testing.assert_array_equal(_valid(got, p), fabs[(slice(p.b.G, p.b.G + p.M),) * p.D])


@pytest.mark.parametrize("name", ["cfg1", "cyl1-m16", "cyl2-m8", "cart3-m8", "cart3-m16"])
@pytest.mark.parametrize("order", [2, 4])
def test_restrict_full_and_rhs_match_oracle(name, order):
    # derivs!'s right-hand side (src/streamer.jl:313-318): restrict_full!(q),
    # then per level fill_ghost!(q, l) and rhs_level!(q1, q, l)
    p = CASES[name]()
    if p.helm and order == 4:
        pytest.skip("order 4 is a Laplacian")
    print(f"case {name} order {order}", flush=True)
    geom = solver.geometry_from_code(p.geom)
    disc = solver.LaplacianDiscretization(p.D, order) if not p.helm else _disc(p)
    ctx, (q, q1, _, _) = _device(p)
    q.upload(p.b)
    solver.restrict_full(q)
    solver.rhs(q1, q, geom, p.bc, disc)
    gq, g1 = p.b.like(), p.b.like()
    q.download(gq)
    q1.download(g1)
    o = Oracle.for_problem(p)
    hq, h1 = p.b.copy(), p.b.like()
    o.restrict_full(hq)
    o.rhs(h1, hq, order=order)
    np.testing.assert_array_equal(_valid(gq, p), _valid(hq, p))
    np.testing.assert_array_equal(_valid(g1, p), _valid(h1, p))
    if order == 4:
        assert not np.array_equal(_valid(g1, p), _valid(gq, p))


@pytest.mark.parametrize("name", ["cfg1", "cyl2-m8", "cart3-m8"])
def test_poisson_step_from_the_charge_matches_oracle(name):
    # derivs!'s Poisson block (src/streamer.jl:309-333) from the net charge q:
    # restrict_full!, the right-hand side loop, then poisson_iter warm-started
    # V-cycles; two steps, q rebuilt before each as netcharge! does
    p = CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx, (u, q1, r, u1) = _device(p)
    q = ctx.field(F64)
    for _ in range(2):
        q.upload(p.b)
        solver.poisson_step(u, q1, r, u1, p.s, geom, p.bc, disc, poisson_iter=2, q=q)
    got = p.b.like()
    u.download(got)
    o = Oracle.for_problem(p)
    uo, ro, u1o = p.b.like(), p.b.like(), p.b.like()
    for _ in range(2):
        qo, q1o = p.b.copy(), p.b.like()
        o.restrict_full(qo)
        o.rhs(q1o, qo, order=2)
        for _ in range(2):
            o.fill_ghost_all(uo)
            o.vcycle(uo, q1o, ro, u1o, p.s, ndown=2, ntop=4, nup=2)
    np.testing.assert_array_equal(_valid(got, p), _valid(uo, p))


def _face_ghosts(a, D, M, G):
    """the depth-1 ghost layer of every face (corners excluded), per block"""
    out = []
    for d in range(D):
        for g in (G - 1, G + M):
            idx = tuple(g if e == d else slice(G, G + M) for e in range(D))
            out.append(a[idx + (slice(None),)])
    return out


@pytest.mark.parametrize("name", ["cfg1", "cyl1-m16", "cart3-m8"])
def test_eabs_ghosts_with_extrapolate_const_match_oracle(name):
    # fill_ghost!(eabs, l, conn, ExtrapolateConst()) for every level after
    # electric_field! (src/streamer.jl:353-355): copy, mirror with +1 and
    # coarse-fine interpolation, then the host copy with its ghosts
    p = CASES[name]()
    print(f"case {name}", flush=True)
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx, (u, b, r, u1) = _device(p)
    b.upload(p.b)
    solver.fill_ghost_all(u, p.bc)
    for _ in range(2):
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    e0 = [0.0, -2.0e6, 1.0e5][:p.D]
    eabs = ctx.field(F64)
    solver.electric_field(u, eabs, p.h, e0, p.bc, G=p.b.G, with_faces=False)
    xc = solver.extrapolate_const(p.D)
    solver.fill_ghost_all(eabs, xc)
    got = p.b.like()
    eabs.download(got, bc=xc)
    host = p.b.like()
    u.download(host)
    o = Oracle.for_problem(p)
    o.fill_ghost_all(host)
    _, fabs = o.electric_field(host, p.h, e0)
    ox = Oracle(p.tree, p.conn, p.M, p.b.G, p.geom, xc, p.helm, p.k2)
    ox.fill_ghost_all(fabs)
    G, M = p.b.G, p.M
    np.testing.assert_array_equal(_valid(got, p), fabs[(slice(G, G + M),) * p.D])
    for k, (x, y) in enumerate(zip(_face_ghosts(got.u, p.D, M, G), _face_ghosts(fabs, p.D, M, G))):
        np.testing.assert_array_equal(x, y, err_msg=f"face {k}")


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8", "cart3-helm"])
def test_zero_start_equals_a_solve_from_zeroed_u(name):
    # SRMG_GUESS_ZERO takes resnorm(0) and the first FP64 defect from b alone;
    # every later step must be bitwise that of a solve from a zeroed u
    p = CASES[name]()
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    out = {}
    for mode in (solver.GUESS_NONE, solver.GUESS_ZERO):
        ctx = solver.Context(p.tree, p.conn, p.M)
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        b.upload(p.b)
        if mode == solver.GUESS_ZERO:
            u.upload(p.b)  # garbage the zero start must ignore
        n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12, guess=mode)
        v, b2 = ctx.field(F64), ctx.field(F64)
        b2.upload(p.b)
        if mode == solver.GUESS_ZERO:
            v.upload(p.b)
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        nm, hm = solver.solve_mixed(v, b2, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-9, max_cycles=12,
                                    inner_cycles=2, guess=mode)
        fields = []
        for f in (u, v):
            host = p.b.like()
            f.download(host)
            fields.append(_valid(host, p).copy())
        out[mode] = (h, hm, fields)
    h0, hm0, f0 = out[solver.GUESS_NONE]
    h1, hm1, f1 = out[solver.GUESS_ZERO]
    assert h1[0] == 1.0 and hm1[0] == 1.0
    np.testing.assert_array_equal(h1[1:], h0[1:])
    np.testing.assert_array_equal(hm1[1:], hm0[1:])
    np.testing.assert_array_equal(f1[0], f0[0])
    np.testing.assert_array_equal(f1[1], f0[1])


@pytest.mark.parametrize("name", ["cfg1", "cart3-m8"])
def test_mixed_zero_start_without_a_step_leaves_u_zero(name):
    # the zero start skips clearing u when the first step writes every cell;
    # a solve that takes no step (tolerance already met) must still return 0
    p = CASES[name]()
    geom, disc = solver.geometry_from_code(p.geom), _disc(p)
    ctx = solver.Context(p.tree, p.conn, p.M)
    v, b = ctx.field(F64), ctx.field(F64)
    e, be, re, ue = (ctx.field(F32) for _ in range(4))
    b.upload(p.b)
    v.upload(p.b)  # garbage
    for tol, cyc in ((2.0, 12), (1e-9, 0)):
        n, h = solver.solve_mixed(v, b, e, be, re, ue, p.s, geom, p.bc, disc, tol=tol, max_cycles=cyc,
                                  inner_cycles=2, guess=solver.GUESS_ZERO)
        assert n == 0
        host = p.b.like()
        v.download(host)
        assert not np.any(_valid(host, p)), (tol, cyc)
