"""srmg_regrid on the device: applydelta!'s field transfer (src/refinement.jl:339-370
with interp! and restrict!, src/streamer.jl:191-221) against the oracle's
restatement on the host, bit for bit in FP64; a context carried across a
regrid solves exactly as a fresh context on the new tree with the same
fields; a regrid that is not one level is refused before anything changes."""

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

import regrid_cases as rc
from oracle import Oracle, applydelta_fields
from srmg import solver
from srmg.blockfield import ScalarBlockField
from srmg.solver import F32, F64, REGRID_INTERP, REGRID_RESTRICT
from srmg.tree import Tree, connectivity, parentcoord, subblocks

pytestmark = pytest.mark.gpu


def _host(p, nb, dtype=np.float64):
    return ScalarBlockField(p.D, p.M, p.b.G, nb, dtype)


def _vs(p):
    return (slice(p.b.G, p.b.G + p.M),) * p.D


def _blocks(tree):
    return [b for layer in tree for _, b in layer.pairs]


@pytest.mark.parametrize("D", [2, 3])
def test_regrid_carries_fields_like_the_oracle(D):
    p = rc.problem(D)
    host = rc.random_field(p)
    ref_in = host.copy()
    Oracle.for_problem(p).fill_ghost_all(ref_in)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, w, z = (ctx.field(F64) for _ in range(3))
    for f in (u, w, z):
        f.upload(host)
    t2, c2, nb2, actions = rc.edited(p.tree, p.tree.nblocks())
    print(f"D={D}: {p.tree.nblocks()} -> {nb2} block ids, {len(actions)} actions", flush=True)
    ctx.regrid(t2, c2, fields=[(u, REGRID_INTERP | REGRID_RESTRICT), (w, REGRID_RESTRICT)], bc=p.bc,
               nblocks=nb2)
    vs = _vs(p)
    want = applydelta_fields(ref_in.u, actions, nb2, p.M, p.b.G)
    want_w = applydelta_fields(ref_in.u, actions, nb2, p.M, p.b.G, interp=False)
    new_ids = {a[1] for a in actions if a[0] == "interp"}
    gu, gw, gz = _host(p, nb2), _host(p, nb2), _host(p, nb2)
    u.download(gu)
    w.download(gw)
    z.download(gz)
    for b in _blocks(t2):
        np.testing.assert_array_equal(gu.u[vs + (b - 1,)], want[vs + (b - 1,)])
        # a field without interpolation starts at zero on new blocks
        if b in new_ids:
            assert np.all(gw.u[vs + (b - 1,)] == 0)
        else:
            np.testing.assert_array_equal(gw.u[vs + (b - 1,)], want_w[vs + (b - 1,)])
        # a field not listed is reallocated at zero
        assert np.all(gz.u[vs + (b - 1,)] == 0)


@pytest.mark.parametrize("D", [2, 3])
def test_regrid_carries_fp32_fields(D):
    # the oracle's restatement in FP32 on ghosts filled in FP64 and rounded:
    # coarse-fine ghosts may differ in the last bit from the device's FP32 fill
    p = rc.problem(D)
    host = rc.random_field(p)
    host.u[...] = host.u.astype(np.float32)
    ref_in = host.copy()
    Oracle.for_problem(p).fill_ghost_all(ref_in)
    ctx = solver.Context(p.tree, p.conn, p.M)
    e = ctx.field(F32)
    e.upload(host)
    t2, c2, nb2, actions = rc.edited(p.tree, p.tree.nblocks())
    ctx.regrid(t2, c2, fields=[(e, REGRID_INTERP | REGRID_RESTRICT)], bc=p.bc, nblocks=nb2)
    want = applydelta_fields(ref_in.u.astype(np.float32), actions, nb2, p.M, p.b.G)
    got = _host(p, nb2, np.float32)
    e.download(got)
    vs = _vs(p)
    scale = np.max(np.abs(host.u))
    for b in _blocks(t2):
        np.testing.assert_allclose(got.u[vs + (b - 1,)], want[vs + (b - 1,)], rtol=0, atol=1e-6 * scale)


@pytest.mark.parametrize("D", [2, 3])
def test_solve_after_regrid_equals_a_fresh_context(D):
    p = rc.problem(D)
    geom = solver.geometry_from_code(p.geom)
    disc = solver.LaplacianDiscretization(p.D, 2)
    ctx = solver.Context(p.tree, p.conn, p.M)
    u, b, r, u1 = (ctx.field(F64) for _ in range(4))
    b.upload(p.b)
    solver.fill_ghost_all(u, p.bc)
    for _ in range(2):
        solver.vcycle(u, b, r, u1, p.s, p.tree, p.conn, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    t2, c2, nb2, _ = rc.edited(p.tree, p.tree.nblocks())
    both = REGRID_INTERP | REGRID_RESTRICT
    ctx.regrid(t2, c2, fields=[(u, both), (b, both)], bc=p.bc, nblocks=nb2)
    hu, hb = _host(p, nb2), _host(p, nb2)
    u.download(hu)
    b.download(hb)
    ctx2 = solver.Context(t2, c2, p.M)
    v, b2, r2, v1 = (ctx2.field(F64) for _ in range(4))
    v.upload(hu)
    b2.upload(hb)
    for uu, bb, rr, ww, cx in ((u, b, r, u1, ctx), (v, b2, r2, v1, ctx2)):
        solver.fill_ghost_all(uu, p.bc)
        for _ in range(2):
            solver.vcycle(uu, bb, rr, ww, p.s, t2, c2, geom, p.bc, disc, ndown=2, ntop=4, nup=2)
    ga, gb = _host(p, nb2), _host(p, nb2)
    u.download(ga)
    v.download(gb)
    vs = _vs(p)
    for blk in _blocks(t2):
        np.testing.assert_array_equal(ga.u[vs + (blk - 1,)], gb.u[vs + (blk - 1,)])
    assert not np.array_equal(ga.u, hu.u)


def test_regrid_refuses_more_than_one_level_and_changes_nothing():
    p = rc.problem(2)
    tree = p.tree
    # levels 1-2 of the tree without the sibling group above some level-3 blocks
    par = next(c for c, _ in tree[1].pairs if not tree.isleaf(2, c))
    gone = set(subblocks(parentcoord(par)))
    small = Tree(2, tree.rootsize, 2)
    for c, blk in tree[0].pairs:
        small[0][c] = blk
    for c, blk in tree[1].pairs:
        if c not in gone:
            small[1][c] = blk
    small.sync()
    cs = connectivity(small)
    ctx = solver.Context(small, cs, p.M, nblocks=tree.nblocks())
    u = ctx.field(F64)
    host = rc.random_field(p)
    u.upload(host)
    with pytest.raises(solver.SrmgError, match="parent is not in the old tree"):
        ctx.regrid(tree, p.conn, fields=[(u, REGRID_INTERP)], bc=p.bc)
    with pytest.raises(solver.SrmgError, match="boundary conditions"):
        ctx.regrid(tree, p.conn, fields=[(u, REGRID_INTERP)])
    # the context still holds the small tree and the field its values
    back = p.b.like()
    u.download(back)
    vs = _vs(p)
    for blk in _blocks(small):
        np.testing.assert_array_equal(back.u[vs + (blk - 1,)], host.u[vs + (blk - 1,)])


def test_forest_regrid_equals_single_tree_regrids():
    # srmg_regrid_forest: each tree of a forest regridded with its own delta
    # gives, tree by tree, the bits of regridding it alone
    probs = [rc.problem(2, seed=s) for s in (1, 2)] + [rc.problem(3, seed=3)]
    probs = [p for p in probs if p.D == 2]  # one dimension per context
    edits = [rc.edited(p.tree, p.tree.nblocks()) for p in probs]
    hosts = [rc.random_field(p, seed=10 + k) for k, p in enumerate(probs)]
    ctx = solver.Context.forest([(p.tree, p.conn) for p in probs], probs[0].M)
    u = ctx.field(F64)
    u.upload(solver.concat_fields(hosts))
    both = REGRID_INTERP | REGRID_RESTRICT
    ctx.set_forest([(t2, c2) for t2, c2, _, _ in edits], fields=[(u, both)], bc=probs[0].bc,
                   nblocks=[nb2 for _, _, nb2, _ in edits])
    nbs = [nb2 for _, _, nb2, _ in edits]
    got = solver.concat_fields([_host(p, nb) for p, nb in zip(probs, nbs)])
    u.download(got)
    parts = solver.split_field(got, nbs)
    vs = _vs(probs[0])
    for k, (p, (t2, c2, nb2, _)) in enumerate(zip(probs, edits)):
        c1 = solver.Context(p.tree, p.conn, p.M)
        v = c1.field(F64)
        v.upload(hosts[k])
        c1.regrid(t2, c2, fields=[(v, both)], bc=p.bc, nblocks=nb2)
        ref = _host(p, nb2)
        v.download(ref)
        for blk in _blocks(t2):
            np.testing.assert_array_equal(parts[k][vs + (blk - 1,)], ref.u[vs + (blk - 1,)])
