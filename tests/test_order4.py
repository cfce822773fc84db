"""The compact 4th-order Laplacian (LaplacianDiscretization{D,4}, Cartesian;
src/laplacian.jl:133-152, 305-340) and its four-colour smoother (FourColors,
src/multigrid.jl:206-232, 268-295) in the oracle, pinned by what the
mathematics fixes: the operator is exact on polynomials (L4 u = h^2 R(lap u)
with R = I + h^2 lap / 12 on quadratics), the four colours partition the cells
and a smoothing pass touches only its colour, and the V-cycle solution
converges faster under refinement than with the order-2 operator."""

import numpy as np
import pytest

from oracle import Oracle
from srmg import configs
from srmg.tree import box_stencil, connectivity


def complete_problem(D, root, M, levels=2, bcd=None, seed=0):
    bcd = bcd or tuple((-1, -1) for _ in range(D))
    return configs.uniform_problem(D, root, M, levels, 0, bcd, seed=seed,
                                   refine=lambda l, c: np.ones(len(c), bool))


def order4_oracle(p):
    return Oracle(p.tree, connectivity(p.tree, box_stencil(p.D)), p.M, p.b.G, 0, p.bc, order=4)


def _centres(p, blk, level):
    """cell-centre coordinates (in level-1 cell units) of every stored cell of a block, ghosts included"""
    G, M, D = p.b.G, p.M, p.D
    c = p.tree[level - 1].coord[blk]
    h = 1.0 / 2 ** (level - 1)
    ax = [((c[d] - 1) * M + np.arange(-G, M + G) + 0.5) * h for d in range(D)]
    return np.meshgrid(*ax, indexing="ij"), h


@pytest.mark.parametrize("D", [2, 3])
def test_order4_operator_is_exact_on_polynomials(D):
    # u = a quartic; L4 u / h^2 = lap u + h^2 lap(lap u) / 12 (R on a quadratic lap u)
    p = complete_problem(D, (2,) * D, 16 if D == 2 else 8)
    o = order4_oracle(p)
    u = p.b.like()
    if D == 2:
        f = lambda x, y: 0.3 * x * x - 0.7 * x * y + 0.2 * y ** 3 + 0.05 * x * x * y * y  # noqa: E731
        lap = lambda x, y: 0.6 + 1.2 * y + 0.1 * (x * x + y * y)  # noqa: E731
        lap2 = 0.4
    else:
        f = lambda x, y, z: 0.3 * x * x - 0.7 * y * z + 0.2 * z ** 3 + 0.1 * x * y * z  # noqa: E731
        lap = lambda x, y, z: 0.6 + 1.2 * z  # noqa: E731
        lap2 = 0.0
    for level in (1, 2):
        for _, blk in p.tree[level - 1].pairs:
            X, h = _centres(p, blk, level)
            u.u[..., blk - 1] = f(*X)
    zero, r = p.b.like(), p.b.like()
    for level in (1, 2):
        o.residual_level(r, u, zero, 1.0, level)  # r = -L4 u
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * D
    for level in (1, 2):
        for _, blk in p.tree[level - 1].pairs:
            X, h = _centres(p, blk, level)
            want = h * h * (lap(*X) + h * h * lap2 / 12.0)
            # cancellation in the stencil sum: rounding relative to the size of u
            np.testing.assert_allclose(-r.u[vs + (blk - 1,)], want[vs], rtol=0, atol=1e-14 * np.max(np.abs(u.u)))


@pytest.mark.parametrize("D", [2, 3])
def test_four_colours_partition_the_cells(D):
    p = complete_problem(D, (2,) * D, 16 if D == 2 else 8, levels=1)
    o = order4_oracle(p)
    rng = np.random.default_rng(D)
    u = p.b.like()
    u.u[...] = rng.standard_normal(u.u.shape)
    b = p.b.like()
    b.u[...] = rng.standard_normal(b.u.shape)
    o.fill_ghost(u, 1)
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * D
    seen = np.zeros(u.u[vs].shape, int)
    for c in range(4):
        before = u.u[vs].copy()
        o.gs4_level(u, b, 1.0, 1.0, 1, c >> 1, c & 1)
        changed = before != u.u[vs]
        seen += changed
        assert changed.sum() == changed.size // 4
        o.fill_ghost(u, 1)
    assert np.all(seen == 1)


@pytest.mark.parametrize("order", [2, 4])
def test_order4_vcycles_converge_faster_than_order2(order):
    # L u = -h^2 R q with q = -lap u* for u* = sin(pi x) sin(pi y) on the unit
    # square (Dirichlet by the odd mirror): per halving of h the error falls by
    # about 4 with the order-2 operator and by 11.5 with the compact order-4
    # one (about h^3.5: the mirror boundary closure is not fourth order)
    errs = []
    for root in (2, 4):
        p = complete_problem(2, (root, root), 8, levels=2)
        o = order4_oracle(p) if order == 4 else Oracle.for_problem(p)
        q = p.b.like()
        ustar = p.b.like()
        n1 = root * 8
        for level in (1, 2):
            for _, blk in p.tree[level - 1].pairs:
                X, h = _centres(p, blk, level)
                x, y = X[0] / n1, X[1] / n1
                q.u[..., blk - 1] = 2 * np.pi ** 2 * np.sin(np.pi * x) * np.sin(np.pi * y)
                ustar.u[..., blk - 1] = np.sin(np.pi * x) * np.sin(np.pi * y)
        q1 = q.like()
        o.rhs(q1, q, order=order)
        u, r, u1 = q.like(), q.like(), q.like()
        s = (1.0 / (2 * n1)) ** 2 * 4  # level-1 spacing squared
        for _ in range(40):
            o.fill_ghost_all(u)
            o.vcycle(u, q1, r, u1, s, ndown=2, ntop=40, nup=2)
        G, M = p.b.G, p.M
        vs = (slice(G, G + M),) * 2
        e = max(np.max(np.abs(u.u[vs + (blk - 1,)] - ustar.u[vs + (blk - 1,)])) for _, blk in p.tree[1].pairs)
        errs.append(e)
    ratio = errs[0] / errs[1]
    print(f"order {order}: errors {errs}, ratio {ratio:.2f}")
    if order == 2:
        assert 3.5 < ratio < 4.5
    else:
        assert 10 < ratio < 17


@pytest.mark.parametrize("geom", [1, 2])
def test_cylindrical_order4_operator_is_exact_on_low_polynomials(geom):
    # LaplacianDiscretization{2,4} x CylindricalGeometry{geom}: L4 u = h^2 lap u
    # (lap in axisymmetric cylindrical coordinates, lap lap u = 0) for r^2, z^2,
    # r^2 z and z^3, with r the coordinate of dimension `geom`
    p = complete_problem(2, (2, 3), 16)
    o = Oracle(p.tree, connectivity(p.tree, box_stencil(2)), p.M, p.b.G, geom, p.bc, order=4)
    cases = [(lambda r, z: r * r, lambda r, z: 4.0 + 0 * r), (lambda r, z: z * z, lambda r, z: 2.0 + 0 * r),
             (lambda r, z: r * r * z, lambda r, z: 4.0 * z), (lambda r, z: z ** 3, lambda r, z: 6.0 * z)]
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * 2
    for f, lap in cases:
        u = p.b.like()
        for level in (1, 2):
            for _, blk in p.tree[level - 1].pairs:
                X, h = _centres(p, blk, level)
                r, z = (X[0], X[1]) if geom == 1 else (X[1], X[0])
                u.u[..., blk - 1] = f(r, z)
        zero, rr = p.b.like(), p.b.like()
        for level in (1, 2):
            o.residual_level(rr, u, zero, 1.0, level)
        for level in (1, 2):
            for _, blk in p.tree[level - 1].pairs:
                X, h = _centres(p, blk, level)
                r, z = (X[0], X[1]) if geom == 1 else (X[1], X[0])
                np.testing.assert_allclose(-rr.u[vs + (blk - 1,)], (h * h * lap(r, z))[vs], rtol=0,
                                           atol=1e-14 * (np.max(np.abs(u.u)) + 1))
