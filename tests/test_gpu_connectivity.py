"""srmg_connectivity: connectivity! (src/connectivity.jl:106-149) on the GPU
against the host restatement (srmg.tree.connectivity), record for record and
in the same order, with the StarStencil and the BoxStencil; trees with
recycled block ids (after applydelta), 2D and 3D, the 3D workload's tree; a
tree whose nesting is broken is refused as the reference asserts."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

import regrid_cases as rc
from srmg import configs, solver
from srmg.tree import Tree, box_stencil, connectivity, parentcoord, subblocks

pytestmark = pytest.mark.gpu


def _same(a, b, L):
    for name in ("neighbor", "boundary", "refboundary", "child"):
        for l in range(L):
            x, y = getattr(a, name)[l], getattr(b, name)[l]
            assert len(x) == len(y), (name, l)
            if len(x):
                np.testing.assert_array_equal(x, y.reshape(x.shape), err_msg=f"{name} level {l + 1}")


def _trees():
    out = {}
    for D in (2, 3):
        p = rc.problem(D)
        out[f"case{D}d"] = p.tree
        out[f"case{D}d-edited"] = rc.edited(p.tree, p.tree.nblocks())[0]
    out["cfg1"] = configs.make_cfg1(0).tree
    return out


@pytest.mark.parametrize("box", [False, True])
def test_device_connectivity_equals_the_host_restatement(box):
    for name, tree in _trees().items():
        print(f"tree {name} ({tree.nblocks()} blocks), box={box}", flush=True)
        st = box_stencil(tree.D) if box else None
        want = connectivity(tree, st)
        got = solver.connectivity_device(tree, solver.BOX_STENCIL if box else solver.STAR_STENCIL)
        _same(got, want, len(tree))


def test_device_connectivity_of_the_3d_workload():
    p = configs.make_cfg4(0)
    got = solver.connectivity_device(p.tree)
    _same(got, p.conn, len(p.tree))


def test_device_connectivity_refuses_broken_nesting():
    p = rc.problem(2)
    tree = p.tree
    # drop the sibling group above some level-3 blocks from level 2
    par = next(c for c, _ in tree[1].pairs if not tree.isleaf(2, c))
    gone = set(subblocks(parentcoord(par)))
    bad = Tree(2, tree.rootsize, 3)
    for li in range(3):
        for c, blk in tree[li].pairs:
            if not (li == 1 and c in gone):
                bad[li][c] = blk
    bad.sync()
    with pytest.raises(solver.SrmgError, match="nesting broken|no parent"):
        solver.connectivity_device(bad)
