"""Host connectivity! (src/connectivity.jl:106-149) with both stencils of
src/stencil.jl: the BoxStencil's order, and the box records against what the
geometry fixes (the star records are the box records with one nonzero
offset; every Neighbor has its reverse; a Boundary's bnd marks the
dimensions that leave the domain)."""

import numpy as np
import pytest

import regrid_cases as rc
from srmg.tree import box_stencil, connectivity, star_stencil


def test_box_stencil_order():
    assert box_stencil(2) == [(-1, -1), (0, -1), (1, -1), (-1, 0), (1, 0), (-1, 1), (0, 1), (1, 1)]
    b3 = box_stencil(3)
    assert len(b3) == 26 and b3[0] == (-1, -1, -1) and b3[1] == (0, -1, -1) and b3[-1] == (1, 1, 1)
    assert (0, 0, 0) not in b3
    assert star_stencil(3) == [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]


@pytest.mark.parametrize("D", [2, 3])
def test_box_records_contain_the_star_records(D):
    p = rc.problem(D)
    star = connectivity(p.tree)
    box = connectivity(p.tree, box_stencil(D))
    for l in range(len(p.tree)):
        for name, w in (("neighbor", 2), ("boundary", 1), ("refboundary", 2)):
            # same records; within a block the two stencils order the faces differently
            b = getattr(box, name)[l]
            keep = np.count_nonzero(b[:, w:w + D], axis=1) == 1 if len(b) else np.zeros(0, bool)
            got = sorted(map(tuple, b[keep].tolist())) if len(b) else []
            assert got == sorted(map(tuple, getattr(star, name)[l].tolist()))
        np.testing.assert_array_equal(box.child[l], star.child[l])
        # every neighbour relation has its reverse
        nb = box.neighbor[l]
        pairs = {(int(r[0]), int(r[1]), tuple(int(x) for x in r[2:])) for r in nb}
        for a, b_, f in pairs:
            assert (b_, a, tuple(-x for x in f)) in pairs
        # bnd marks exactly the dimensions that leave the level's domain
        layer = p.tree[l]
        for r in box.boundary[l]:
            c = layer.coord[int(r[0])]
            s, bnd = r[1:1 + D], r[1 + D:1 + 2 * D]
            for d in range(D):
                q = c[d] + s[d]
                out = not (layer.domain[d][0] <= q <= layer.domain[d][1])
                assert bnd[d] == (s[d] if out else 0)
            assert any(bnd) and r[1 + 2 * D] == l + 1
