"""Regrid on the host side and the oracle of applydelta!'s field part
(src/refinement.jl:339-370, src/interp.jl:23-60, src/restrict.jl:5-26):
the tree edit against the reference's rules, and the per-block interp! and
restrict! of the oracle against what the mathematics fixes (linear fields are
interpolated exactly, a restriction is the mean over the children)."""

import numpy as np
import pytest

import regrid_cases as rc
from oracle import Oracle, applydelta_fields, interp_block, restrict_block
from srmg.tree import REFINE, REMOVE, applydelta, connectivity, parentcoord, subblocks, subcoord


@pytest.mark.parametrize("D", [2, 3])
def test_applydelta_edits_the_tree_like_the_reference(D):
    p = rc.problem(D)
    tree = p.tree
    nb = tree.nblocks()
    delta = rc.delta_for(tree)
    removed = [b for b, d in delta.items() if d == REMOVE]
    refined = [b for b, d in delta.items() if d == REFINE]
    coord3 = {b: c for c, b in tree[2].pairs}
    coord2 = {b: c for c, b in tree[1].pairs}
    t2, conn2, nb2, actions = rc.edited(tree, nb)
    # removed leaves are gone, refined leaves have their 2^D children
    for b in removed:
        assert not t2[2].hasblock(coord3[b])
    for b in refined:
        assert all(t2[2].hasblock(ch) for ch in subblocks(coord2[b]))
    # finest level first: every restriction precedes every interpolation
    kinds = [a[0] for a in actions]
    assert kinds == ["restrict"] * len(removed) + ["interp"] * (len(refined) << D)
    # freed ids are popped last-in first-out, then new ids follow nb
    freed = [a[2] for a in actions if a[0] == "restrict"]
    new_ids = [a[1] for a in actions if a[0] == "interp"]
    assert new_ids[:len(freed)] == freed[::-1]
    assert new_ids[len(freed):] == list(range(nb + 1, nb2 + 1))
    assert nb2 == nb + (len(refined) << D) - len(removed)
    # actions carry the parent and the subblock of each block
    for kind, a, b, sub in actions:
        c = coord3[b] if kind == "restrict" else t2[2].coord[a]
        assert sub == subcoord(c)
        par = t2[1].get(parentcoord(c))
        assert par == (a if kind == "restrict" else b)
    # still properly nested: connectivity! finds every coarse-fine parent
    assert sum(len(x) for x in conn2.refboundary) > 0


def test_applydelta_skips_level_one_non_leaves_and_bounds():
    p = rc.problem(2)
    tree = p.tree
    import copy
    t2 = copy.deepcopy(tree)
    # a level-1 leaf, a non-leaf and a finest-level leaf marked REFINE, and a
    # level-2 leaf marked REMOVE with minlevel 2: nothing changes
    l1_leaf = next(b for c, b in tree[0].pairs if tree.isleaf(1, c))
    l2_parent = next(b for c, b in tree[1].pairs if not tree.isleaf(2, c))
    l3_leaf = tree[2].pairs[0][1]
    l2_leaf = next(b for c, b in tree[1].pairs if tree.isleaf(2, c))
    delta = {l1_leaf: REFINE, l2_parent: REMOVE, l3_leaf: REFINE, l2_leaf: REMOVE}
    nb2, actions = applydelta(t2, delta, [], tree.nblocks(), minlevel=2)
    assert actions == [] and nb2 == tree.nblocks()
    assert [sorted(l.pairs) for l in t2] == [sorted(l.pairs) for l in tree]


@pytest.mark.parametrize("D", [2, 3])
def test_interp_block_is_exact_on_linear_fields(D):
    M, G = (16, 2) if D == 2 else (8, 2)
    rng = np.random.default_rng(D)
    a0, g = rng.standard_normal(), rng.standard_normal(D)
    # parent cell centres at integer + 1/2 (spacing 1), ghosts included
    x = np.meshgrid(*[np.arange(M + 2 * G) - G + 0.5] * D, indexing="ij")
    src = a0 + sum(g[d] * x[d] for d in range(D))
    for sub in [(1,) * D, (2,) * D, (1, 2) + (1,) * (D - 2)]:
        dest = np.zeros_like(src)
        interp_block(dest, src, sub, M, G)
        # fine cell centres: spacing 1/2 from the subblock's corner
        xf = np.meshgrid(*[(sub[d] - 1) * M / 2 + (np.arange(M) + 0.5) / 2 for d in range(D)], indexing="ij")
        want = a0 + sum(g[d] * xf[d] for d in range(D))
        np.testing.assert_allclose(dest[(slice(G, G + M),) * D], want, rtol=0, atol=1e-12)


@pytest.mark.parametrize("D", [2, 3])
def test_restrict_block_is_the_mean_of_the_children(D):
    M, G = (16, 2) if D == 2 else (8, 2)
    rng = np.random.default_rng(10 + D)
    src = rng.standard_normal((M + 2 * G,) * D)
    dest = np.full_like(src, 7.0)
    sub = (2, 1) + (2,) * (D - 2)
    restrict_block(dest, sub, src, M, G)
    v = src[(slice(G, G + M),) * D]
    mean = v.reshape(sum([[M // 2, 2] for _ in range(D)], [])).mean(axis=tuple(range(1, 2 * D, 2)))
    h = M // 2
    sl = tuple(slice(G + (sub[d] - 1) * h, G + sub[d] * h) for d in range(D))
    np.testing.assert_allclose(dest[sl], mean, rtol=1e-14, atol=1e-15)
    # the rest of the block is untouched
    mask = np.ones(dest.shape, bool)
    mask[sl] = False
    assert np.all(dest[mask] == 7.0)
    # a constant is restricted and interpolated exactly
    c = np.full_like(src, 3.25)
    d2 = np.zeros_like(src)
    interp_block(d2, c, sub, M, G)
    assert np.all(d2[(slice(G, G + M),) * D] == 3.25)


def test_applydelta_fields_keeps_ids_and_applies_in_order():
    p = rc.problem(2)
    host = rc.random_field(p)
    o = Oracle.for_problem(p)
    o.fill_ghost_all(host)
    t2, conn2, nb2, actions = rc.edited(p.tree, p.tree.nblocks())
    out = applydelta_fields(host.u, actions, nb2, p.M, host.G)
    vs = (slice(host.G, host.G + p.M),) * 2
    kept = [b for layer in t2 for _, b in layer.pairs
            if b <= p.tree.nblocks() and b not in {a[1] for a in actions if a[0] == "interp"}]
    for b in kept:
        parent_of_removed = any(a[0] == "restrict" and a[1] == b for a in actions)
        if not parent_of_removed:
            assert np.array_equal(out[vs + (b - 1,)], host.u[vs + (b - 1,)])
    # without interpolation a new block keeps what its id held (a recycled id:
    # the removed block's values, as newblock!-free reuse leaves them) or zero
    # (a new id); without restriction parents keep their values
    out2 = applydelta_fields(host.u, actions, nb2, p.M, host.G, interp=False, restrict=False)
    for kind, a, b, sub in actions:
        if kind == "interp":
            want = host.u[vs + (a - 1,)] if a <= p.tree.nblocks() else 0.0
            assert np.all(out2[vs + (a - 1,)] == want)
        else:
            assert np.array_equal(out2[vs + (a - 1,)], host.u[vs + (a - 1,)])
