"""Regrid cases shared by the CPU and GPU regrid tests: a three-level tree and
a refinement delta as refdelta! would leave it (src/refinement.jl:220-331):
one complete sibling group of finest-level leaves marked REMOVE (their parent
not refined) and two level-2 leaves whose in-domain star neighbours all exist
marked REFINE, so the edited tree stays properly nested."""

import copy

import numpy as np

from srmg import configs
from srmg.tree import REFINE, REMOVE, applydelta, connectivity, star_stencil, subblocks


def _crit(D):
    if D == 2:
        def crit(l, c):
            if l == 1:
                return (c[:, 0] >= 2) & (c[:, 1] >= 2)
            return (c[:, 0] >= 5) & (c[:, 1] >= 5) & (c[:, 1] <= 6)
        return crit

    def crit3(l, c):
        if l == 1:
            return (c[:, 0] >= 2) & (c[:, 1] <= 2)
        return (c[:, 0] >= 5) & (c[:, 1] <= 2) & (c[:, 2] <= 4)
    return crit3


def problem(D, seed=0):
    if D == 2:
        return configs.uniform_problem(2, (3, 4), 16, 3, 1, ((1, -1), (-1, -1)), refine=_crit(2), seed=seed)
    return configs.uniform_problem(3, (3, 2, 2), 8, 3, 0, ((1, 1), (1, 1), (-1, -1)), refine=_crit(3), seed=seed)


def delta_for(tree):
    """REMOVE for the 2^D children of one level-2 block, REFINE for two
    level-2 leaves with every in-domain star neighbour present."""
    D = tree.D
    l2, l3 = tree[1], tree[2]
    parents = [c for c, _ in l2.pairs if not tree.isleaf(2, c)]
    assert len(parents) >= 2, "the case needs two refined level-2 blocks"
    delta = {}
    for ch in subblocks(parents[0]):
        delta[l3.get(ch)] = REMOVE
    picked = []
    for c, blk in l2.pairs:
        if not tree.isleaf(2, c):
            continue
        ok = True
        for s in star_stencil(D):
            q = tuple(a + b for a, b in zip(c, s))
            if l2.in_domain(q) and not l2.hasblock(q):
                ok = False
        if ok:
            picked.append(blk)
        if len(picked) == 2:
            break
    assert len(picked) == 2, "the case needs two refinable level-2 leaves"
    for blk in picked:
        delta[blk] = REFINE
    return delta


def edited(tree, nblocks):
    """(new tree, its connectivity, new id count, actions) after applydelta."""
    t2 = copy.deepcopy(tree)
    nb2, actions = applydelta(t2, delta_for(tree), [], nblocks)
    return t2, connectivity(t2), nb2, actions


def random_field(p, seed=1):
    f = p.b.like()
    f.u[...] = np.random.default_rng(seed).standard_normal(f.u.shape)
    return f
