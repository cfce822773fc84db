"""rhs_level! (src/multigrid.jl:363-388) in the oracle: the order-2 forms are
the identity; the compact 4th-order R of LaplacianDiscretization{D,4}
(src/laplacian.jl:142-149, 167-175, 187-195, 322-337) is I + L2/12, L2 being
the order-2 operator at unit spacing (its weights sum to 12 with the
radial weights w- + w+ = 2), checked against the oracle's own residual;
restrict_full! against a level-by-level mean."""

import numpy as np
import pytest

from oracle import Oracle
from srmg import configs


def _case(D, geom):
    if D == 2:
        bcd = ((1, -1), (-1, -1))
        return configs.uniform_problem(2, (3, 4), 16, 3, geom, bcd, seed=5,
                                       refine=lambda l, c: (c[:, 0] <= 2) & (c[:, 1] >= 2) & (c[:, 1] <= 3 * l))
    bcd = ((1, 1), (1, 1), (-1, -1))
    return configs.uniform_problem(3, (2, 3, 2), 8, 3, 0, bcd, seed=6,
                                   refine=lambda l, c: (c[:, 0] == l) & (c[:, 1] >= 2))


def _vs(p):
    return (slice(p.b.G, p.b.G + p.M),) * p.D


@pytest.mark.parametrize("D,geom", [(2, 0), (2, 1), (2, 2), (3, 0)])
def test_order4_rhs_is_identity_plus_laplacian_over_12(D, geom):
    p = _case(D, geom)
    o = Oracle.for_problem(p)
    q = p.b.copy()
    q.u[...] = np.random.default_rng(D + geom).standard_normal(q.u.shape)
    b1 = q.like()
    o.rhs(b1, q, order=4)
    # -L2 q from the oracle's residual with b = 0 and s = 1 (ghosts as filled above)
    zero, r = q.like(), q.like()
    for l in range(1, len(p.tree) + 1):
        o.residual_level(r, q, zero, 1.0, l)
    vs = _vs(p)
    want = q.u[vs] - r.u[vs] / 12.0
    np.testing.assert_allclose(b1.u[vs], want, rtol=0, atol=1e-14 * np.max(np.abs(q.u)))
    # order 2: the identity on valid cells
    b2 = q.like()
    o.rhs(b2, q, order=2)
    assert np.array_equal(b2.u[vs], q.u[vs])


def test_restrict_full_is_the_mean_level_by_level():
    p = _case(2, 0)
    o = Oracle.for_problem(p)
    u = p.b.copy()
    u.u[...] = np.random.default_rng(1).standard_normal(u.u.shape)
    want = u.copy()
    o.restrict_full(u)
    # finest first: level-3 children into level 2, then level 2 (updated) into level 1
    G, M, h = p.b.G, p.M, p.M // 2
    for li in (2, 1):
        for c, blk in p.tree[li].pairs:
            par = p.tree[li - 1].get(tuple((x - 1) // 2 + 1 for x in c))
            sub = tuple((x - 1) % 2 for x in c)
            v = want.u[G:G + M, G:G + M, blk - 1]
            mean = (v[0::2, 0::2] + v[1::2, 0::2] + v[0::2, 1::2] + v[1::2, 1::2]) / 4
            want.u[G + sub[0] * h:G + sub[0] * h + h, G + sub[1] * h:G + sub[1] * h + h, par - 1] = mean
    vs = _vs(p)
    np.testing.assert_allclose(u.u[vs], want.u[vs], rtol=0, atol=1e-15)
