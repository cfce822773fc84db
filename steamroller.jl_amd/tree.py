"""Host-side block tree and connectivity.

The tree stays on the host, as in the reference: a list of per-level block
layers mapping block coordinates to block ids, and per-level relation lists
(boundary, neighbour, refinement boundary, child) that the device library
flattens into its own tables at every regrid.

Reference interfaces mirrored here:
  * ``BlockLayer{D}`` / ``sync!`` / ``get`` / ``hasblock``  -- src/blocklayer.jl:9-115
  * ``Tree(D, domain, levels)`` / ``populate!`` / ``isleaf`` -- src/tree.jl:13-44, 105-115
  * ``parentcoord`` / ``subcoord`` / ``subblocks``         -- src/mesh.jl:40-61
  * ``StarStencil`` / ``BoxStencil`` iteration order        -- src/stencil.jl:16-56
  * ``connectivity!`` and the four edge records              -- src/connectivity.jl:6-42, 106-149
  * proper-nesting closure used when refining               -- src/refinement.jl:286-331
  * the tree edit of ``applydelta!``                         -- src/refinement.jl:339-370

Coordinates and block ids are 1-based, exactly as in the reference, so the
int64 records produced here are bit-for-bit the isbits structs a Julia caller
would hand over through ``ccall`` (``Neighbor{D}`` = from, to, face[D]; etc.).
"""

from __future__ import annotations

import itertools
from typing import Callable, Dict, Iterable, List, Sequence, Tuple

import numpy as np

Coord = Tuple[int, ...]


def parentcoord(c: Coord) -> Coord:
    """src/mesh.jl:48-50: ``div(c[d] - 1, 2) + 1``."""
    return tuple((x - 1) // 2 + 1 for x in c)


def subcoord(c: Coord) -> Coord:
    """src/mesh.jl:59-61: ``rem(c[d] - 1, 2) + 1`` (entries are 1 or 2)."""
    return tuple((x - 1) % 2 + 1 for x in c)


def subblocks(b: Coord) -> List[Coord]:
    """src/mesh.jl:40-42: the 2^D children of block ``b``, column-major order."""
    D = len(b)
    ranges = [range(2 * (x - 1) + 1, 2 * x + 1) for x in b]
    # column-major (first index fastest), like CartesianIndices
    return [tuple(reversed(t)) for t in itertools.product(*reversed(ranges))]


def star_stencil(D: int) -> List[Coord]:
    """src/stencil.jl:16-33: (-1 in dim 1), (+1 in dim 1), (-1 in dim 2), ..."""
    out = []
    for d in range(D):
        for v in (-1, 1):
            s = [0] * D
            s[d] = v
            out.append(tuple(s))
    return out


def box_stencil(D: int) -> List[Coord]:
    """src/stencil.jl:36-56: every offset in -1..1 (first index fastest) but zero."""
    out = []
    for o in column_major([range(-1, 2)] * D):
        if any(o):
            out.append(o)
    return out


def column_major(ranges: Sequence[range]) -> Iterable[Coord]:
    """Iterate the cartesian product with the first index fastest."""
    for t in itertools.product(*reversed(ranges)):
        yield tuple(reversed(t))


class BlockLayer:
    """One refinement level (src/blocklayer.jl:9-24).

    ``domain`` is the per-dimension inclusive 1-based block range at this
    level, ``rootsize * 2^(level-1)`` (src/tree.jl:13-15).
    """

    def __init__(self, level: int, domain: Sequence[Tuple[int, int]]):
        self.level = level
        self.domain = tuple(domain)
        self.index: Dict[Coord, int] = {}
        self.coord: Dict[int, Coord] = {}
        self.pairs: List[Tuple[Coord, int]] = []

    def in_domain(self, c: Coord) -> bool:
        return all(lo <= x <= hi for x, (lo, hi) in zip(c, self.domain))

    def __setitem__(self, c: Coord, blk: int):
        self.index[c] = blk
        self.coord[blk] = c

    def get(self, c: Coord, default: int = 0) -> int:
        return self.index.get(c, default)

    def hasblock(self, c: Coord) -> bool:
        return c in self.index

    def sync(self):
        """src/blocklayer.jl:48-53. Pairs are kept sorted by block id so the
        iteration order (and hence every record list) is deterministic."""
        self.pairs = sorted(self.index.items(), key=lambda kv: kv[1])

    def __len__(self):
        return len(self.pairs)

    def __iter__(self):
        return iter(self.pairs)


class Tree(list):
    """``Tree{D} = Vector{BlockLayer{D}}`` (src/tree.jl:6-15)."""

    def __init__(self, D: int, rootsize: Sequence[int], levels: int):
        super().__init__()
        self.D = D
        self.rootsize = tuple(rootsize)
        for l in range(1, levels + 1):
            f = 1 << (l - 1)
            self.append(BlockLayer(l, [(1, r * f) for r in rootsize]))

    def nblocks(self) -> int:
        return sum(len(layer) for layer in self)

    def sync(self):
        for layer in self:
            layer.sync()

    def isleaf(self, lvl: int, c: Coord) -> bool:
        """src/tree.jl:105-115 (``lvl`` is 1-based)."""
        if lvl >= len(self):
            return True
        nxt = self[lvl]  # level lvl + 1
        return not any(nxt.hasblock(ch) for ch in subblocks(c))

    def finestlevel(self) -> int:
        """src/tree.jl:141: last non-empty level (1-based)."""
        last = 0
        for i, layer in enumerate(self):
            if len(layer.index):
                last = i + 1
        return last

    def level_of_block(self) -> Dict[int, int]:
        return {blk: layer.level for layer in self for _, blk in layer.pairs}


def populate(tree: Tree, level: int | None = None) -> int:
    """src/tree.jl:31-44: fill every level up to ``level`` completely."""
    level = len(tree) if level is None else level
    i = 0
    for l in range(1, level + 1):
        layer = tree[l - 1]
        for c in column_major([range(lo, hi + 1) for lo, hi in layer.domain]):
            i += 1
            layer[c] = i
    tree.sync()
    return i


def refined_tree(D: int, rootsize: Sequence[int], levels: int,
                 crit: Callable[[int, np.ndarray], np.ndarray]) -> Tree:
    """Build a properly nested tree by refining where ``crit`` asks.

    ``crit(level, coords)`` receives an (n, D) int array of 1-based block
    coordinates at ``level`` and returns a boolean mask of blocks to refine.
    Level 1 is complete. After marking, the nesting closure of
    src/refinement.jl:286-331 is applied: a refined block needs every in-domain
    star neighbour to exist at its level (so every child has the parents of
    its star neighbours, the condition asserted at src/connectivity.jl:129);
    a missing neighbour is created by refining its parent, recursively.
    Block ids are assigned level by level in column-major coordinate order,
    as ``populate!`` would (src/tree.jl:31-44).
    """
    tree = Tree(D, rootsize, levels)
    stencil = star_stencil(D)
    exists: List[set] = [set() for _ in range(levels)]
    refined: List[set] = [set() for _ in range(levels)]
    exists[0] = set(column_major([range(lo, hi + 1) for lo, hi in tree[0].domain]))

    def require(l: int, q: Coord):
        # make block q exist at level l (1-based)
        if q in exists[l - 1]:
            return
        assert l > 1, "level 1 is complete"
        require_refined(l - 1, parentcoord(q))

    def require_refined(l: int, p: Coord):
        if p in refined[l - 1]:
            return
        require(l, p)
        refined[l - 1].add(p)
        for ch in subblocks(p):
            exists[l].add(ch)
        layer = tree[l - 1]
        for s in stencil:
            q = tuple(a + b for a, b in zip(p, s))
            if layer.in_domain(q):
                require(l, q)

    for l in range(1, levels):
        cand = sorted(exists[l - 1])
        if not cand:
            break
        arr = np.array(cand, dtype=np.int64).reshape(-1, D)
        mask = np.asarray(crit(l, arr), dtype=bool)
        for c, m in zip(cand, mask):
            if m:
                require_refined(l, c)

    blk = 0
    for l in range(1, levels + 1):
        layer = tree[l - 1]
        for c in sorted(exists[l - 1], key=lambda t: tuple(reversed(t))):
            blk += 1
            layer[c] = blk
    tree.sync()
    return tree


# RefDelta (src/refinement.jl:207)
REMOVE, KEEP, REFINE = -1, 0, 1


def applydelta(tree: Tree, delta: Dict[int, int], freeblocks: List[int], nblocks: int, minlevel: int = 1,
               maxlevel: int | None = None) -> Tuple[int, List[Tuple[str, int, int, Coord]]]:
    """The tree edit of applydelta! (src/refinement.jl:339-370), in place.

    Blocks are visited as ``@blocks order=serial`` does (finest level first,
    each level in its ``pairs`` order); level-1 blocks and blocks that are not
    leaves are skipped. A leaf marked REMOVE above ``minlevel`` is deleted and
    its id pushed on ``freeblocks``; a leaf marked REFINE below ``maxlevel``
    gets its 2^D children (``subblocks`` order), each taking an id popped from
    ``freeblocks`` or, when none is free, the next new id (``newblock!``).
    ``delta`` maps block ids to REMOVE / KEEP / REFINE (missing ids: KEEP).

    Returns the new number of block ids (the host fields' length) and the
    field actions in the reference's order: ``("restrict", parent, blk, sub)``
    for a removed block and ``("interp", new, blk, sub)`` for a new child, sub
    being the 1-based ``subcoord``. The device regrid (``srmg_regrid``) needs
    only the edited tree; the actions drive the host restatement in tests.
    """
    maxlevel = len(tree) if maxlevel is None else min(maxlevel, len(tree))
    actions: List[Tuple[str, int, int, Coord]] = []
    for lvl in range(len(tree), 0, -1):
        layer = tree[lvl - 1]
        for c, blk in list(layer.pairs):
            if lvl <= 1 or not tree.isleaf(lvl, c):
                continue
            d = delta.get(blk, KEEP)
            if d == REMOVE and lvl > minlevel:
                pblk = tree[lvl - 2].get(parentcoord(c))
                actions.append(("restrict", pblk, blk, subcoord(c)))
                del layer.index[c]
                del layer.coord[blk]
                freeblocks.append(blk)
            elif d == REFINE and lvl < maxlevel:
                for ch in subblocks(c):
                    if freeblocks:
                        nblk = freeblocks.pop()
                    else:
                        nblocks += 1
                        nblk = nblocks
                    tree[lvl][ch] = nblk
                    actions.append(("interp", nblk, blk, subcoord(ch)))
    tree.sync()
    return nblocks, actions


class Connectivity:
    """Per-level relation lists (src/connectivity.jl:51-71).

    Each list is an int64 array of records, laid out as the reference's
    isbits structs (src/connectivity.jl:6-42):
      boundary    (block, face[D], bnd[D], level)     -> 2 + 2D words
      neighbor    (from, to, face[D])                 -> 2 + D words
      refboundary (fine, coarse, face[D], subblock[D]) -> 2 + 2D words
      child       (fine, coarse, subblock[D])          -> 2 + D words
    """

    def __init__(self, D: int, levels: int):
        self.D = D
        self.boundary = [np.zeros((0, 2 + 2 * D), np.int64) for _ in range(levels)]
        self.neighbor = [np.zeros((0, 2 + D), np.int64) for _ in range(levels)]
        self.refboundary = [np.zeros((0, 2 + 2 * D), np.int64) for _ in range(levels)]
        self.child = [np.zeros((0, 2 + D), np.int64) for _ in range(levels)]


def connectivity(tree: Tree, stencil: Sequence[Coord] | None = None) -> Connectivity:
    """src/connectivity.jl:106-149; ``stencil`` defaults to the StarStencil
    (order-2 operators); :func:`box_stencil` gives the BoxStencil."""
    D = tree.D
    conn = Connectivity(D, len(tree))
    stencil = star_stencil(D) if stencil is None else list(stencil)
    for li, layer in enumerate(tree):
        bnd, nbr, ref, chd = [], [], [], []
        for c, blk in layer.pairs:
            for s in stencil:
                q = tuple(a + b for a, b in zip(c, s))
                if not layer.in_domain(q):
                    bvec = tuple(0 if layer.domain[d][0] <= q[d] <= layer.domain[d][1]
                                 else s[d] for d in range(D))
                    bnd.append((blk, *s, *bvec, layer.level))
                    continue
                other = layer.get(q)
                if other != 0:
                    nbr.append((blk, other, *s))
                elif layer.level > 1:
                    pblk = tree[li - 1].get(parentcoord(q))
                    assert pblk != 0, "Proper nesting broken"
                    ref.append((blk, pblk, *s, *subcoord(q)))
            if layer.level > 1:
                pblk = tree[li - 1].get(parentcoord(c))
                assert pblk != 0, "Malformed tree"
                chd.append((blk, pblk, *subcoord(c)))
        if bnd:
            conn.boundary[li] = np.array(bnd, np.int64)
        if nbr:
            conn.neighbor[li] = np.array(nbr, np.int64)
        if ref:
            conn.refboundary[li] = np.array(ref, np.int64)
        if chd:
            conn.child[li] = np.array(chd, np.int64)
    return conn


def flatten_pairs(tree: Tree) -> Tuple[np.ndarray, np.ndarray]:
    """All levels' ``pairs`` as (coord[D], blk) int64 records plus level offsets."""
    D = tree.D
    parts, off = [np.zeros((0, D + 1), np.int64)], [0]
    for layer in tree:
        n = len(layer.pairs)
        flat = itertools.chain.from_iterable(c + (blk,) for c, blk in layer.pairs)
        parts.append(np.fromiter(flat, np.int64, count=n * (D + 1)).reshape(n, D + 1))
        off.append(off[-1] + n)
    return np.ascontiguousarray(np.concatenate(parts)), np.array(off, np.int64)


def flatten_records(lists: List[np.ndarray], width: int) -> Tuple[np.ndarray, np.ndarray]:
    """Concatenate per-level record arrays; return (records, level offsets)."""
    off = [0]
    for a in lists:
        off.append(off[-1] + len(a))
    if off[-1]:
        arr = np.concatenate([a.reshape(-1, width) for a in lists]).astype(np.int64)
    else:
        arr = np.zeros((0, width), np.int64)
    return np.ascontiguousarray(arr), np.array(off, np.int64)


def boundaryconditions(desc: Sequence[Tuple[int, int]]) -> np.ndarray:
    """src/boundary.jl:28-38: Int8 3^D table, +1 Neumann / -1 Dirichlet.

    Returned flat in Julia column-major order: entry for face ``f`` is at
    ``sum((f[d] + 1) * 3**d)``.
    """
    D = len(desc)
    a = np.ones((3,) * D, dtype=np.int8)
    for i, (lo, hi) in enumerate(desc):
        sl_lo = [slice(None)] * D
        sl_hi = [slice(None)] * D
        sl_lo[i] = slice(0, 1)
        sl_hi[i] = slice(2, 3)
        a[tuple(sl_lo)] *= lo
        a[tuple(sl_hi)] *= hi
    return np.ascontiguousarray(a.ravel(order="F"))


def leaf_cells(tree: Tree, M: int) -> int:
    """Number of valid cells in leaf blocks (the composite unknowns)."""
    n = 0
    for layer in tree:
        for c, _ in layer.pairs:
            if tree.isleaf(layer.level, c):
                n += M ** tree.D
    return n
