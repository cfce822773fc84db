"""Pins for the CPU oracle (oracle/srmg_oracle.c).

The reference ships no tests, golden vectors or fixtures for this path, so the
oracle is pinned by what the mathematics fixes: an independently assembled
sparse operator solved with scipy, exactness of the transfer operators on
linear data, the mirror boundary signs, the fixed point of the V-cycle on
composite AMR operators, and a pure-Python transliteration of the reference's
loops on tiny trees (tests/test_transliteration.py).
"""

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from assemble import grid_operator, level_to_grid, grid_to_level
from oracle import Oracle
from srmg import configs
from srmg.blockfield import ScalarBlockField
from srmg.tree import boundaryconditions, connectivity, refined_tree

CASES = [
    # D, rootsize, M, geom, bc
    (2, (2, 3), 8, 0, ((-1, -1), (-1, -1))),
    (2, (2, 4), 8, 1, ((1, -1), (-1, -1))),
    (2, (3, 2), 8, 2, ((-1, 1), (1, -1))),
    (3, (2, 2, 2), 4, 0, ((1, 1), (1, 1), (-1, -1))),
    (3, (2, 1, 2), 4, 0, ((-1, 1), (1, -1), (-1, -1))),
]


def _problem(D, rootsize, M, geom, bc, levels=1, helm=False, k2=0.0, seed=0, refine=None):
    return configs.uniform_problem(D, rootsize, M, levels, geom, bc, helm=helm, k2=k2,
                                   seed=seed, refine=refine)


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("helm", [False, True])
def test_coarse_solve_matches_spsolve(case, helm):
    D, rootsize, M, geom, bc = case
    k2 = 3.0e4 if helm else 0.0
    p = _problem(D, rootsize, M, geom, bc, helm=helm, k2=k2)
    o = Oracle.for_problem(p)
    u = p.b.like()
    assert o.coarse_solve(u, p.b, p.s) == 0
    bvec, N = level_to_grid(p.b, p.tree, 1, M)
    A = grid_operator(N, geom, p.bc, -(p.s * k2) if helm else 0.0)
    ref = spla.spsolve(A, -p.s * bvec)
    got, _ = level_to_grid(u, p.tree, 1, M)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("case", CASES[:4])
def test_vcycle_on_complete_two_levels_matches_fine_grid(case):
    """With level 2 complete the composite problem is the uniform level-2 grid."""
    D, rootsize, M, geom, bc = case
    p = _problem(D, rootsize, M, geom, bc, levels=2, refine=lambda l, c: np.ones(len(c), bool))
    o = Oracle.for_problem(p)
    u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
    o.fill_ghost_all(u)
    for _ in range(40):
        o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
    bvec, N = level_to_grid(p.b, p.tree, 2, M)
    A = grid_operator(N, geom, p.bc)
    ref = spla.spsolve(A, -(p.s / 4) * bvec)
    got, _ = level_to_grid(u, p.tree, 2, M)
    assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


@pytest.mark.parametrize("case", CASES)
def test_red_black_half_sweeps_match_matrix_splitting(case):
    """One gauss_seidel_iter! sweep = update of the cells with odd global index
    sum from the old values (the RedBlack{0} set), then of the even ones, each
    u_i -= omega (A u + s b)_i / c with c = diagelm = -2D, the stencil centre
    (src/multigrid.jl:172-177): at boundary cells the mirror ghost enters A u,
    not the divisor."""
    D, rootsize, M, geom, bc = case
    p = _problem(D, rootsize, M, geom, bc, seed=3)
    o = Oracle.for_problem(p)
    rng = np.random.default_rng(1)
    u = p.b.like()
    for _, blk in p.tree[0].pairs:
        u.valid(blk)[...] = rng.standard_normal((M,) * D)
    o.fill_ghost_all(u)
    uvec, N = level_to_grid(u, p.tree, 1, M)
    bvec, _ = level_to_grid(p.b, p.tree, 1, M)
    A = grid_operator(N, geom, p.bc).tocsr()
    omega = 0.9
    idx = np.indices(N).reshape(D, -1, order="F")
    jsum = (idx + 1).sum(axis=0)
    x = uvec.copy()
    for parity in (1, 0):
        sel = (jsum % 2) == parity
        upd = x - omega * (A @ x + p.s * bvec) / (-2.0 * D)
        x[sel] = upd[sel]
    o.gs_iter(u, p.b, omega, p.s, 1, 1)
    got, _ = level_to_grid(u, p.tree, 1, M)
    assert np.max(np.abs(got - x)) <= 1e-12 * np.max(np.abs(x))
    # the first half-sweep alone touches exactly the odd cells
    u2 = p.b.like()
    grid_to_level(uvec, N, u2, p.tree, 1, M)
    o.fill_ghost_all(u2)
    o.gs_level(u2, p.b, omega, p.s, 1, 0)
    got2, _ = level_to_grid(u2, p.tree, 1, M)
    changed = got2 != uvec
    assert np.all(changed[(jsum % 2) == 1] | (uvec[(jsum % 2) == 1] == got2[(jsum % 2) == 1]))
    assert not np.any(changed[(jsum % 2) == 0])


def _two_level_patch(D, M, rootsize):
    """A level-2 patch strictly inside the domain (one refined root block plus
    the nesting closure), for transfer-operator checks away from boundaries."""
    centre = tuple((r + 1) // 2 for r in rootsize)

    def crit(l, c):
        return np.all(c == np.array(centre), axis=1)
    tree = refined_tree(D, rootsize, 2, crit)
    return tree, connectivity(tree)


@pytest.mark.parametrize("D,M,rootsize", [(2, 8, (4, 4)), (3, 4, (4, 4, 4))])
def test_interp_ghosts_exact_for_linear_data(D, M, rootsize):
    tree, conn = _two_level_patch(D, M, rootsize)
    G = 2
    f = ScalarBlockField(D, M, G, tree.nblocks())
    g = np.array([0.37, -1.3, 0.71][:D])

    def lin(level, blk, c):
        # value at cell centres in level-1 cell units, ghosts included
        hl = 1.0 / (1 << (level - 1))
        ax = [hl * (M * (c[d] - 1) + np.arange(-G, M + G) + 0.5) for d in range(D)]
        grids = np.meshgrid(*ax, indexing="ij")
        return 2.5 + sum(g[d] * grids[d] for d in range(D))
    for c, blk in tree[0].pairs:
        f.block(blk)[...] = lin(1, blk, c)
    o = Oracle(tree, conn, M, G, 0, boundaryconditions(((1, 1),) * D))
    o.fill_ghost_interp(f, 2)
    assert len(conn.refboundary[1]) > 0
    for rec in conn.refboundary[1]:
        fine = int(rec[0])
        face = rec[2:2 + D]
        c = tree[1].coord[fine]
        want = lin(2, fine, c)
        sl = []
        for d in range(D):
            if face[d] == 0:
                sl.append(slice(G, G + M))
            elif face[d] < 0:
                sl.append(slice(0, G))
            else:
                sl.append(slice(M + G, M + 2 * G))
        sl = tuple(sl)
        np.testing.assert_allclose(f.block(fine)[sl], want[sl], rtol=0, atol=1e-13)


@pytest.mark.parametrize("D,M,rootsize", [(2, 8, (4, 4)), (3, 4, (4, 4, 4))])
def test_restrict_and_interp_add_on_linear_data(D, M, rootsize):
    tree, conn = _two_level_patch(D, M, rootsize)
    G = 2
    g = np.array([0.37, -1.3, 0.71][:D])

    def lin(level, c):
        hl = 1.0 / (1 << (level - 1))
        ax = [hl * (M * (c[d] - 1) + np.arange(-G, M + G) + 0.5) for d in range(D)]
        grids = np.meshgrid(*ax, indexing="ij")
        return -0.75 + sum(g[d] * grids[d] for d in range(D))
    o = Oracle(tree, conn, M, G, 0, boundaryconditions(((1, 1),) * D))
    # restriction of fine linear data = coarse linear data (mean of children)
    f = ScalarBlockField(D, M, G, tree.nblocks())
    for c, blk in tree[1].pairs:
        f.block(blk)[...] = lin(2, c)
    o.restrict_level(f, 2)
    vs = (slice(G, G + M),) * D
    for rec in conn.child[1]:
        coarse, sb = int(rec[1]), rec[2:2 + D]
        cc = tree[0].coord[coarse]
        want = lin(1, cc)[vs]
        sl = tuple(slice((int(sb[d]) - 1) * M // 2, int(sb[d]) * M // 2) for d in range(D))
        np.testing.assert_allclose(f.valid(coarse)[sl], want[sl], rtol=0, atol=1e-13)
    # interpolation of coarse linear data added onto zero fine data is exact
    u = ScalarBlockField(D, M, G, tree.nblocks())
    u1 = ScalarBlockField(D, M, G, tree.nblocks())
    for c, blk in tree[0].pairs:
        u1.block(blk)[...] = lin(1, c)
    o.interp_add_level(u, u1, 2)
    for c, blk in tree[1].pairs:
        np.testing.assert_allclose(u.valid(blk), lin(2, c)[vs], rtol=0, atol=1e-13)


@pytest.mark.parametrize("D", [2, 3])
def test_boundary_mirror_and_copy(D):
    M, G = 4, 2
    rootsize = (2,) * D
    tree = refined_tree(D, rootsize, 1, lambda l, c: np.zeros(len(c), bool))
    conn = connectivity(tree)
    bcd = tuple((-1, 1) if d % 2 == 0 else (1, -1) for d in range(D))
    bc = boundaryconditions(bcd)
    o = Oracle(tree, conn, M, G, 0, bc)
    f = ScalarBlockField(D, M, G, tree.nblocks())
    rng = np.random.default_rng(0)
    # a global array of level-1 cells; copy ghosts must reproduce neighbours
    N = [M * r for r in rootsize]
    glob = rng.standard_normal(N)
    for c, blk in tree[0].pairs:
        sl = tuple(slice(M * (c[d] - 1), M * c[d]) for d in range(D))
        f.valid(blk)[...] = glob[sl]
    o.fill_ghost(f, 1)
    # reconstruct the expected padded global array with mirror ghosts (depth G)
    pad = np.pad(glob, G)
    for d in range(D):
        lo_sign, hi_sign = bcd[d]
        for k in range(G):
            src_lo = [slice(None)] * D
            dst_lo = [slice(None)] * D
            dst_lo[d] = G - 1 - k
            src_lo[d] = G + k
            pad[tuple(dst_lo)] = lo_sign * pad[tuple(src_lo)]
            dst_hi = [slice(None)] * D
            src_hi = [slice(None)] * D
            dst_hi[d] = G + N[d] + k
            src_hi[d] = G + N[d] - 1 - k
            pad[tuple(dst_hi)] = hi_sign * pad[tuple(src_hi)]
    for c, blk in tree[0].pairs:
        for d in range(D):
            for side in (0, 1):
                # face slab of depth G (valid range in the other dims)
                sl_blk, sl_pad = [], []
                for e in range(D):
                    if e == d:
                        if side == 0:
                            sl_blk.append(slice(0, G))
                            sl_pad.append(slice(M * (c[e] - 1), M * (c[e] - 1) + G))
                        else:
                            sl_blk.append(slice(M + G, M + 2 * G))
                            sl_pad.append(slice(M * c[e] + G, M * c[e] + 2 * G))
                    else:
                        sl_blk.append(slice(G, G + M))
                        sl_pad.append(slice(M * (c[e] - 1) + G, M * c[e] + G))
                np.testing.assert_array_equal(f.block(blk)[tuple(sl_blk)], pad[tuple(sl_pad)])


def _tiny_amr(D, geom, bc, seed=0, M=8):
    rootsize = (3, 4) if D == 2 else (2, 2, 2)

    def crit(l, c):
        if l == 1:
            return np.all(c == c[:1], axis=1) if D == 3 else (c[:, 0] == 1) & (c[:, 1] == 2)
        return (c[:, 0] == 1) & (c[:, 1] == 3) if D == 2 else np.all(c == np.array([2, 2, 2]), axis=1)
    levels = 3 if D == 2 else 2
    return configs.uniform_problem(D, rootsize, M if D == 2 else 4, levels, geom, bc, seed=seed,
                                   refine=crit)


def _composite_operator(o, p):
    """Leaf residual A(x) of leaf values x: covered cells restricted, ghosts
    filled level by level, residual formed on every leaf block."""
    tree, M, D = p.tree, p.M, p.D
    leaves = [(l, c, blk) for l, layer in enumerate(tree, 1) for c, blk in layer.pairs
              if tree.isleaf(l, c)]
    n = len(leaves) * M ** D

    def apply(x):
        u = p.b.like()
        for k, (l, c, blk) in enumerate(leaves):
            u.valid(blk)[...] = x[k * M ** D:(k + 1) * M ** D].reshape((M,) * D, order="F")
        for l in range(len(tree), 1, -1):
            o.restrict_level(u, l)
        o.fill_ghost_all(u)
        r = p.b.like()
        for l in range(1, len(tree) + 1):
            o.residual_level(r, u, p.b, p.s / 4 ** (l - 1), l)
        return np.concatenate([r.valid(blk).ravel(order="F") for (l, c, blk) in leaves])
    return apply, n, leaves


@pytest.mark.parametrize("D,geom,bcd", [(2, 1, ((1, -1), (-1, -1))), (2, 0, ((-1, -1), (1, -1))),
                                        (3, 0, ((1, 1), (1, 1), (-1, -1)))])
def test_vcycle_fixed_point_is_the_composite_solution(D, geom, bcd):
    p = _tiny_amr(D, geom, bcd)
    o = Oracle.for_problem(p)
    apply, n, leaves = _composite_operator(o, p)
    a0 = apply(np.zeros(n))
    K = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        K[:, i] = apply(e) - a0
    x = np.linalg.solve(K, -a0)
    u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
    o.fill_ghost_all(u)
    for _ in range(60):
        o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
    M = p.M
    got = np.concatenate([u.valid(blk).ravel(order="F") for (l, c, blk) in leaves])
    assert np.max(np.abs(got - x)) <= 1e-9 * np.max(np.abs(x))
    # one more cycle leaves the converged solution unchanged
    before = u.u.copy()
    o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
    vs = (slice(p.b.G, p.b.G + M),) * D
    assert np.max(np.abs(u.u[vs] - before[vs])) <= 1e-12 * np.max(np.abs(before[vs]))


@pytest.mark.parametrize("D,geom,bcd", [(2, 1, ((1, -1), (-1, -1))), (3, 0, ((1, 1), (1, 1), (-1, -1)))])
def test_resnorm_is_weighted_leaf_sum(D, geom, bcd):
    p = _tiny_amr(D, geom, bcd, seed=5)
    o = Oracle.for_problem(p)
    u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
    o.fill_ghost_all(u)
    o.vcycle(u, b, r, u1, p.s, ndown=1, ntop=2, nup=1)
    rr = p.b.like()
    tot = 0.0
    for l in range(1, len(p.tree) + 1):
        o.residual_level(rr, u, b, p.s / 4 ** (l - 1), l)
        for c, blk in p.tree[l - 1].pairs:
            if p.tree.isleaf(l, c):
                tot += np.sum(rr.valid(blk) ** 2) * 2.0 ** (-D * (l - 1))
    assert o.resnorm(r, b, u, p.s) == pytest.approx(np.sqrt(tot), rel=1e-13)


def test_vcycle_converges_on_cfg1_with_exact_coarse_solve():
    p = configs.make_cfg1(0)
    o = Oracle.for_problem(p)
    u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
    r0 = o.resnorm(r, b, u, p.s)
    o.fill_ghost_all(u)
    rel = []
    for _ in range(10):
        o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=4, nup=2, coarse_exact=True)
        rel.append(o.resnorm(r, b, u, p.s) / r0)
    assert rel[-1] < 1e-8
    assert all(b2 < a2 for a2, b2 in zip(rel[:-1], rel[1:]))


@pytest.mark.parametrize("D,M,rootsize", [(2, 8, (4, 4)), (3, 4, (4, 4, 4))])
def test_electric_field_of_a_linear_potential(D, M, rootsize):
    # u = g . x with dyadic h and integer g: every difference and quotient is
    # exact, so on faces away from the domain boundary E = -g + e0 exactly
    # (coarse-fine ghosts by linear interpolation are exact for linear data),
    # and |E| = sqrt(sum (2E)^2 / 4) at cells whose faces are all interior
    tree, conn = _two_level_patch(D, M, rootsize)
    G = 2
    h = 0.25
    g = np.array([3.0, -5.0, 2.0][:D])
    e0 = np.array([0.5, -2.0, 1.0][:D])
    u = ScalarBlockField(D, M, G, tree.nblocks())
    for level in (1, 2):
        hl = h / (1 << (level - 1))
        for c, blk in tree[level - 1].pairs:
            ax = [hl * (M * (c[d] - 1) + np.arange(-G, M + G) + 0.5) for d in range(D)]
            grids = np.meshgrid(*ax, indexing="ij")
            u.block(blk)[...] = sum(g[d] * grids[d] for d in range(D))
    o = Oracle(tree, conn, M, G, 0, boundaryconditions(((1, 1),) * D))
    o.fill_ghost_all(u)
    f, fabs = o.electric_field(u, h, e0)
    S = M + 2 * G + 1
    want_abs = np.sqrt(sum((2 * (e0[d] - g[d])) ** 2 / 4 for d in range(D)))
    nface = ncell = 0
    for level in (1, 2):
        ext = [r << (level - 1) for r in rootsize]
        for c, blk in tree[level - 1].pairs:
            fb = f[..., blk - 1]
            for d in range(D):
                sl = [slice(G, G + M)] * D
                lo = G + (1 if c[d] == 1 else 0)
                hi = G + M + (0 if c[d] == ext[d] else 1)
                sl[d] = slice(lo, hi)
                vals = fb[tuple(sl) + (d,)]
                assert np.all(vals == e0[d] - g[d]), (level, c, d)
                nface += vals.size
            inner = [slice(G + (1 if c[d] == 1 else 0), G + M - (1 if c[d] == ext[d] else 0)) for d in range(D)]
            cells = fabs[tuple(inner) + (blk - 1,)]
            assert np.all(cells == want_abs)
            ncell += cells.size
    assert nface > 0 and ncell > 0
    # faces outside the valid range are not written
    assert np.all(f[(slice(0, G),) + (slice(None),) * D] == 0)


AXES = [(8, False, -1, -1), (16, False, 1, -1), (64, True, 1, -1), (128, False, -1, 1),
        (32, True, 1, 1), (64, False, 1, 1), (256, True, 1, -1), (512, False, -1, -1)]


@pytest.mark.parametrize("n,radial,alow,ahigh", AXES)
def test_ql_axis_decomposition_pinned_by_jacobi_and_dense_eigenvalues(n, radial, alow, ahigh):
    """The exact coarse solve's implicit QL eigensolver against two independent
    routes: the cyclic Jacobi method and numpy's dense eigenvalues of the
    assembled finite-volume operator; both factorizations reconstruct T."""
    from assemble import axis_matrix
    from oracle import axis_decomp
    T = axis_matrix(n, radial, alow, ahigh).toarray()
    lam, P, Q = axis_decomp(n, radial, alow, ahigh, "ql")
    lj, Pj, Qj = axis_decomp(n, radial, alow, ahigh, "jacobi") if n <= 128 else (None, None, None)
    scale = np.max(np.abs(T))
    np.testing.assert_allclose(P @ np.diag(lam) @ Q, T, rtol=0, atol=1e-12 * scale)
    np.testing.assert_allclose(P @ Q, np.eye(n), rtol=0, atol=1e-12)
    np.testing.assert_allclose(np.sort(lam), np.sort(np.linalg.eigvals(T).real), rtol=0, atol=1e-12 * scale)
    if lj is not None:
        np.testing.assert_allclose(np.sort(lam), np.sort(lj), rtol=0, atol=1e-13 * scale)
        np.testing.assert_allclose(Pj @ np.diag(lj) @ Qj, T, rtol=0, atol=1e-12 * scale)


@pytest.mark.parametrize("n,radial,alow,ahigh", AXES)
def test_library_axis_decomposition_equals_the_oracle_bitwise(n, radial, alow, ahigh):
    """The library's host eigensolver (srmg_axis_decomp, which the exact coarse
    solve uploads) and the oracle's produce the same bits, so exact-coarse
    V-cycles can agree bit for bit on the GPU (no device work here)."""
    from oracle import axis_decomp
    from srmg import solver
    lam, P, Q = solver.axis_decomp(n, radial, alow, ahigh)
    ol, oP, oQ = axis_decomp(n, radial, alow, ahigh, "ql")
    np.testing.assert_array_equal(lam, ol)
    np.testing.assert_array_equal(P, oP)
    np.testing.assert_array_equal(Q, oQ)
