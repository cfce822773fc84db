"""Free boundary conditions (src/freebc.jl) on the CPU: the oracle's restatement
against scipy's sine transforms, a pure-Python transliteration of the
reference's loops, the boundary kernel by two routes, and the physics the
construction is for — an open radial boundary."""

import math

import numpy as np
import pytest

from oracle import Oracle, dst_matrix
from srmg.blockfield import ScalarBlockField
from srmg.configs import fill_field
from srmg.freebc import boundary_kernel
from srmg.tree import boundaryconditions, connectivity, refined_tree

M, G = 16, 2
DIM = 1  # CylindricalGeometry{1}: dimension 1 is the radius


def _outer_refined_tree():
    """3 levels on a 3 x 4 root; levels 2 and 3 reach the outer radial
    boundary, so fine outer blocks have coarse-fine faces along z."""
    def crit(l, c):
        if l == 1:
            return (c[:, 0] >= 2) & (c[:, 1] >= 2) & (c[:, 1] <= 3)
        return (c[:, 0] == 6) & (c[:, 1] == 5)
    tree = refined_tree(2, (3, 4), 3, crit)
    return tree, connectivity(tree)


# ---- transliteration of src/freebc.jl (pure Python, the reference's loops) ----

def _setouter_py(q, b, tree, conn, dim, level, s):
    D = 2
    other = dim % D + 1
    face0 = tuple(1 if d == dim - 1 else 0 for d in range(D))
    for l in range(1, level + 1):
        sl = s / (1 << (l - 1)) ** 2
        p = 2 ** (level - l)
        layer = tree[l - 1]
        for rec in conn.boundary[l - 1]:
            blk, face = int(rec[0]), tuple(int(x) for x in rec[1:1 + D])
            if face != face0:
                continue
            B = layer.coord[blk]
            i0 = (B[other - 1] - 1) * M
            for i in range(1, M + 1):
                I = [0, 0]
                I[dim - 1], I[other - 1] = G + M, G + i
                J = [0, 0]
                J[dim - 1], J[other - 1] = M * (B[dim - 1] - 1) + M, M * (B[other - 1] - 1) + i
                bmean = sum(float(x) for x in b[p * (i0 + i - 1): p * (i0 + i)]) / p
                factor = abs(1 + 1 / (2 * J[dim - 1] - 1))
                q[I[0] - 1, I[1] - 1, blk - 1] += 2 * factor * bmean / sl
        k = layer.domain[dim - 1][1]
        for rec in conn.refboundary[l - 1]:
            iblk = int(rec[0])
            face = tuple(int(x) for x in rec[2:2 + D])
            B = layer.coord[iblk]
            if B[dim - 1] != k:
                continue
            dr = face[other - 1]
            i = G + 1 if dr == -1 else M + G
            I = [0, 0]
            I[dim - 1], I[other - 1] = G + M, i
            j = (B[other - 1] - 1) * M + i - G
            J = [0, 0]
            J[dim - 1], J[other - 1] = M * (B[dim - 1] - 1) + M, j
            J1 = [(J[d] - 1) // 2 + 1 + (dr if d == other - 1 else 0) for d in range(D)]
            if dr == 1:
                bmean = sum(float(x) for x in b[p * j: p * (j + 2)]) / (2 * p)
            else:
                bmean = sum(float(x) for x in b[p * (j - 3) - 1: p * (j - 1) - 1]) / (2 * p)
            factor = abs(1 + 1 / (2 * J1[dim - 1] - 1))
            q[I[0] - 1, I[1] - 1, iblk - 1] += 2 * (1 / 4) * factor * bmean / sl


def _fill_ghost_free_py(u, b, tree, dim, level, l):
    other = dim % 2 + 1
    layer = tree[l - 1]
    k = layer.domain[dim - 1][1]
    p = 2 ** (level - l)
    lo, hi = layer.domain[other - 1]
    for iblk in range(lo, hi + 1):
        B = [0, 0]
        B[dim - 1], B[other - 1] = k, iblk
        blk = layer.get(tuple(B))
        if blk == 0:
            continue
        i0 = (iblk - 1) * M
        for i in range(1, M + 1):
            I = [0, 0]
            I[dim - 1], I[other - 1] = G + M + 1, G + i
            bmean = sum(float(x) for x in b[p * (i0 + i - 1): p * (i0 + i)]) / p
            u[I[0] - 1, I[1] - 1, blk - 1] += 2 * bmean


def _oracle(tree, conn):
    return Oracle(tree, conn, M, G, 1, boundaryconditions(((1, -1), (-1, -1))))


# ---- tests ----

@pytest.mark.parametrize("n", [1, 2, 7, 64, 256])
@pytest.mark.parametrize("kind", [2, 3])
def test_sine_matrices_are_fftw_r2r(n, kind):
    """RODFT10 / RODFT01 (src/freebc.jl:45-46) are scipy's DST types 2 / 3."""
    from scipy.fft import dst
    x = np.random.default_rng(n).standard_normal(n)
    y = dst_matrix(n, kind) @ x
    np.testing.assert_allclose(y, dst(x, type=kind), rtol=0, atol=1e-13 * np.max(np.abs(x)) * n)


def test_boundary_kernel_two_routes_and_limit():
    """The kernel from the scaled Bessel functions equals the one from the
    unscaled ones where those do not overflow, and tends to 1/(H k) for large
    k R (I1/I0 -> 1 - 1/(2x), K1/K0 -> 1 + 1/(2x))."""
    from scipy.special import iv, kv
    n, H, R = 512, 2e-2, 5e-3
    f = boundary_kernel(n, H, R)
    k = (math.pi / H) * np.arange(1, n + 1)
    x = k * R
    ok = x < 600
    f2 = (2 / H) / (k[ok] * (iv(1, x[ok]) / iv(0, x[ok]) + kv(1, x[ok]) / kv(0, x[ok])))
    np.testing.assert_allclose(f[ok], f2, rtol=1e-12)
    big = x > 200
    np.testing.assert_allclose(f[big] * H * k[big], 1.0, rtol=2e-5)
    assert np.all(f > 0) and np.all(np.diff(f) < 0)


def test_getouter_takes_the_finest_outer_cell():
    tree, conn = _outer_refined_tree()
    level = len(tree)
    h = 1.0 / 48
    # u = the z coordinate of the cell centre
    u = fill_field(tree, conn, M, h, G, lambda r, z: z)
    o = _oracle(tree, conn)
    n = 4 * 2 ** (level - 1) * M
    u1 = o.freebc_getouter(u, DIM, level, h, (3, 4), n)
    # expected: z of the finest outermost cell covering entry j, over h_l
    hf = h / 2 ** (level - 1)
    for j in range(n):
        zf = (j + 0.5) * hf
        for l in range(level, 0, -1):
            hl = h / 2 ** (l - 1)
            layer = tree[l - 1]
            kz = int(zf // (M * hl)) + 1
            if layer.get((layer.domain[0][1], kz)):
                zc = (math.floor(zf / hl) + 0.5) * hl
                assert u1[j] == pytest.approx(zc / hl, rel=1e-14), (j, l)
                break
        else:
            raise AssertionError(f"entry {j} not covered")


@pytest.mark.parametrize("refined", [False, True])
def test_setouter_and_ghosts_match_the_transliteration(refined):
    if refined:
        tree, conn = _outer_refined_tree()
    else:
        tree = refined_tree(2, (3, 4), 1, lambda l, c: np.zeros(len(c), bool))
        conn = connectivity(tree)
    level = len(tree)
    n = 4 * 2 ** (level - 1) * M
    rng = np.random.default_rng(3)
    b = rng.standard_normal(n)
    q = fill_field(tree, conn, M, 1 / 48, G, lambda r, z: np.sin(7 * r) * np.cos(3 * z))
    qo, qp = q.copy(), q.copy()
    o = _oracle(tree, conn)
    o.freebc_setouter(qo, b, DIM, level, 0.37, (3, 4))
    _setouter_py(qp.u, b, tree, conn, DIM, level, 0.37)
    assert np.array_equal(qo.u, qp.u)
    assert not np.array_equal(qo.u, q.u)
    if refined:
        # a corner correction happened (an outer fine block with a coarse-fine face)
        assert any(len(r) for r in conn.refboundary)
    for l in range(1, level + 1):
        o.fill_ghost_free(qo, b, DIM, level, l, (3, 4))
        _fill_ghost_free_py(qp.u, b, tree, DIM, level, l)
    assert np.array_equal(qo.u, qp.u)


def _uniform(rootsize, h, func):
    tree = refined_tree(2, rootsize, 1, lambda l, c: np.zeros(len(c), bool))
    conn = connectivity(tree)
    return tree, conn, fill_field(tree, conn, M, h, G, func)


def _solve(o, tree, q, s):
    u = q.like()
    r, u1 = q.like(), q.like()
    b = q.copy()
    o.fill_ghost_all(u)
    o.vcycle(u, b, r, u1, s, ndown=2, ntop=4, nup=2, coarse_exact=True)
    return u


def test_free_boundary_approximates_the_open_domain():
    """A charge near the axis in a cylinder r < R: the potential with the free
    boundary at R matches, inside R, the potential of a domain four times wider
    (nearly open) much better than the homogeneous Dirichlet condition does."""
    h = 1.0 / 64
    charge = lambda r, z: np.exp(-(r ** 2 + (z - 1.0) ** 2) / 0.05)  # noqa: E731
    bc = boundaryconditions(((1, -1), (-1, -1)))
    s = h * h
    # the cylinder: R = 0.5, H = 2
    tree, conn, q = _uniform((2, 8), h, charge)
    o = Oracle(tree, conn, M, G, 1, bc)
    uh = _solve(o, tree, q, s)
    f = boundary_kernel(8 * M, 8 * M * h, 2 * M * h)
    q1 = q.copy()
    o.freebc(q1, uh, h, s, DIM, 1, (2, 8), f)
    ufree = _solve(o, tree, q1, s)
    # the wide domain: R = 2
    treew, connw, qw = _uniform((8, 8), h, charge)
    ow = Oracle(treew, connw, M, G, 1, bc)
    uw = _solve(ow, treew, qw, s)
    vs = (slice(G, G + M), slice(G, G + M))

    def inner(u, tr):
        # valid cells of the blocks with r-block <= 2 (r < 0.5), keyed by coordinate
        return {c: u.u[vs + (blk - 1,)] for c, blk in tr[0].pairs if c[0] <= 2}
    ref = inner(uw, treew)
    err = lambda u: max(np.max(np.abs(inner(u, tree)[c] - ref[c])) for c in ref)  # noqa: E731
    scale = max(np.max(np.abs(v)) for v in ref.values())
    e_dir, e_free = err(uh) / scale, err(ufree) / scale
    print(f"max error inside R, relative: Dirichlet {e_dir:.3e}, free {e_free:.3e}")
    assert e_free < 0.1 * e_dir
