"""Seeded synthetic problems for the multigrid path.

Each generator returns a :class:`Problem`: a properly nested block tree, its
connectivity, the operator description and the right-hand side ``b`` in the
reference's host layout. ``b`` is built the way the reference's Poisson caller
builds ``q1`` (src/streamer.jl:309-318): the charge density is evaluated at the
cell centres of every block at every level, then ``restrict_full!`` overwrites
covered coarse cells with the mean of their children; for the order-2 operators
``rhs_level!`` is the identity.

Charge shapes follow the reference's Gaussian initial condition
``A exp(-((r^2 + (z - z0)^2) / w^2))`` (src/initcond.jl:52-60). The scale
``s = h^2 e / eps0`` uses the reference's constants (src/constants.jl:55-56)
and ``s = h^2`` for Helmholtz photoionization terms (src/photo.jl:55).
"""

from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence

import numpy as np

from .blockfield import ScalarBlockField, cell_centers, restrict_full_host
from .tree import (Connectivity, Tree, boundaryconditions, connectivity,
                   leaf_cells, refined_tree)

ELEMENTARY_CHARGE = 1.6021766208e-19   # src/constants.jl:55
EPSILON_0 = 8.854187817620389e-12      # src/constants.jl:56
TORR = 133.32236842105263              # src/constants.jl:139
CENTI = 0.01
MILLI = 0.001
BAR = 100000.0

GEOM_CARTESIAN = 0
GEOM_CYL1 = 1   # CylindricalGeometry{1}: dimension 1 is r
GEOM_CYL2 = 2   # CylindricalGeometry{2}


@dataclass
class Problem:
    name: str
    D: int
    M: int
    h: float
    tree: Tree
    conn: Connectivity
    geom: int
    bc: np.ndarray          # flat int8 3^D table (column-major)
    s: float
    b: ScalarBlockField     # right-hand side, host layout
    k2: float = 0.0         # Helmholtz k^2 (0 and helm=False for Poisson)
    helm: bool = False
    seed: int = 0
    extra: dict = field(default_factory=dict)

    @property
    def n_leaf(self) -> int:
        return leaf_cells(self.tree, self.M)

    @property
    def n_stored(self) -> int:
        return self.tree.nblocks() * self.M ** self.D


def _block_centers(tree: Tree, M: int, h: float, level: int):
    """Cell-centre coordinate arrays for every block of a level, shape
    (nblk, M, ..., M) per dimension, first index = dim-1 position."""
    layer = tree[level - 1]
    coords = np.array([c for c, _ in layer.pairs], np.int64).reshape(-1, tree.D)
    blks = np.array([b for _, b in layer.pairs], np.int64)
    hl = h / (1 << (level - 1))
    D = tree.D
    grids = []
    for d in range(D):
        base = hl * (M * (coords[:, d] - 1))
        off = hl * np.arange(1, M + 1) - hl / 2
        shape = [len(blks)] + [1] * D
        shape[1 + d] = M
        g = (base[:, None] + off[None, :]).reshape([len(blks)] + [M if k == d else 1 for k in range(D)])
        grids.append(np.broadcast_to(g, (len(blks),) + (M,) * D))
    return blks, grids


def fill_field(tree: Tree, conn: Connectivity, M: int, h: float, G: int,
               func: Callable, dtype=np.float64) -> ScalarBlockField:
    """Evaluate ``func(*coords)`` at every cell centre of every block, then
    restrict covered cells (src/streamer.jl:309-313)."""
    f = ScalarBlockField(tree.D, M, G, tree.nblocks(), dtype)
    vs = (slice(G, G + M),) * tree.D
    for l in range(1, len(tree) + 1):
        if len(tree[l - 1]) == 0:
            continue
        blks, grids = _block_centers(tree, M, h, l)
        vals = func(*grids)  # (nblk, M, ..., M), index 1 = dim 1
        # to Fortran block layout (M, ..., M, nblk)
        f.u[vs + (blks - 1,)] = np.moveaxis(vals, 0, -1)
    restrict_full_host(f, tree, conn)
    return f


def gaussian_rz(A, w, z0):
    """src/initcond.jl:52-60 with s=1, extend=0 (cylindrical r, z)."""
    def f(r, z):
        return A * np.exp(-((r ** 2 + (z - z0) ** 2) / w ** 2))
    return f


# --------------------------------------------------------------------------
# cfg1: 2D axisymmetric Gaussian ion cloud, ~20k cells, 3 levels
# --------------------------------------------------------------------------

def make_cfg1(seed: int = 0, G: int = 2) -> Problem:
    """Plain FP64 Poisson, 2D (r, z): r in [0, 5 mm], z in [0, 10 mm], root
    4x8 blocks of 16x16 cells (h = 78.125 um), two refinement levels around a
    Gaussian ion cloud on the axis (w = 0.5 mm, peak 1e20 m^-3, centre
    z_c = 5 mm + U(-0.5, 0.5) mm). BC ((1,-1),(-1,-1)): Neumann on the axis,
    Dirichlet at outer r and both electrodes (src/simulate.jl:244-247); the
    applied 20 kV is the uniform background field, outside the solve."""
    rng = np.random.default_rng(seed)
    D, M = 2, 16
    L = 5e-3
    rootsize = (4, 8)
    h = L / (rootsize[0] * M)
    zc = 5e-3 + rng.uniform(-0.5e-3, 0.5e-3)
    w = 0.5e-3

    def crit(level, coords):
        hl = h / (1 << (level - 1))
        bl = M * hl
        rmin = (coords[:, 0] - 1) * bl
        zlo = (coords[:, 1] - 1) * bl
        zhi = coords[:, 1] * bl
        zdist = np.maximum(0, np.maximum(zlo - zc, zc - zhi))
        return (rmin < 1e-9) & (zdist < (2.5e-3 if level == 1 else 1.5e-3))

    tree = refined_tree(D, rootsize, 3, crit)
    conn = connectivity(tree)
    b = fill_field(tree, conn, M, h, G, gaussian_rz(1e20, w, zc))
    bc = boundaryconditions(((1, -1), (-1, -1)))
    s = h ** 2 * (ELEMENTARY_CHARGE / EPSILON_0)
    return Problem("cfg1", D, M, h, tree, conn, GEOM_CYL1, bc, s, b, seed=seed,
                   extra={"zc": zc, "w": w})


# --------------------------------------------------------------------------
# cfg2: 2D axisymmetric streamer-like tree, 6 levels, ~2M cells
# --------------------------------------------------------------------------

def streamer_charge(z0, length, radius=0.2e-3, shell=20e-6, A_ch=1e19, A_head=2e20):
    """Net charge of a positive streamer: a weakly charged channel of the given
    radius along the axis from ``z0`` to ``z0 + length`` plus a thin surface
    charge shell (Gaussian thickness ``shell``) around its hemispherical head."""
    zh = z0 + length

    def f(r, z):
        along = np.exp(-(np.maximum(0, z0 - z) ** 2 + np.maximum(0, z - zh) ** 2) / radius ** 2)
        ch = A_ch * np.exp(-(r / radius) ** 2) * along
        rho = np.sqrt(r ** 2 + np.maximum(0, z - zh) ** 2)
        # shell hugs the channel side and the head cap
        hd = A_head * np.exp(-((rho - radius) / shell) ** 2) * (z > zh - 4 * radius)
        return ch + hd
    return f


def make_cfg2(seed: int = 0, G: int = 2, target_cells: float = 2.0e6) -> Problem:
    """2D (r, z) streamer-like AMR: root 8x16 blocks of 16x16 cells
    (128x256 level-1 cells), 6 levels, r in [0, 1 cm], z in [0, 2 cm].
    Channel radius 0.2 mm, length 4 mm with seeded axial start, head shell of
    Gaussian thickness 20 um; refined by distance to the channel surface."""
    rng = np.random.default_rng(seed)
    D, M = 2, 16
    rootsize = (8, 16)
    h = 1e-2 / (rootsize[0] * M)
    z0 = 8e-3 + rng.uniform(-1e-3, 1e-3)
    length, radius = 4e-3, 0.2e-3
    zh = z0 + length
    # refinement distance per level, scaled to hit the target size
    base = [3.0e-3, 2.2e-3, 1.5e-3, 1.1e-3, 0.95e-3]

    def crit(level, coords):
        hl = h / (1 << (level - 1))
        bl = M * hl
        rlo, rhi = (coords[:, 0] - 1) * bl, coords[:, 0] * bl
        zlo, zhi = (coords[:, 1] - 1) * bl, coords[:, 1] * bl
        # distance from block box to the channel surface segment r=radius, z in [z0, zh]
        dr = np.maximum(0, np.maximum(rlo - radius, radius - rhi))
        dz = np.maximum(0, np.maximum(zlo - zh, z0 - zlo - bl))
        d = np.sqrt(dr ** 2 + dz ** 2)
        # and to the head cap centre
        dh = np.sqrt(np.maximum(0, rlo) ** 2 + np.maximum(0, np.maximum(zlo - zh, zh - zhi)) ** 2)
        return (d < base[level - 1]) | (dh < 1.5 * base[level - 1])

    tree = refined_tree(D, rootsize, 6, crit)
    conn = connectivity(tree)
    b = fill_field(tree, conn, M, h, G, streamer_charge(z0, length, radius))
    bc = boundaryconditions(((1, -1), (-1, -1)))
    s = h ** 2 * (ELEMENTARY_CHARGE / EPSILON_0)
    return Problem("cfg2", D, M, h, tree, conn, GEOM_CYL1, bc, s, b, seed=seed,
                   extra={"z0": z0})


# --------------------------------------------------------------------------
# cfg4: 3D Cartesian, 8^3 blocks, 5 levels, ~30M cells, branching filaments
# --------------------------------------------------------------------------

def filament_segments(rng, n_fil=3, L=1e-2):
    """Three seeded branching filaments: each a polyline from near the top
    electrode downwards, with one side branch, radius U(0.1, 0.3) mm and
    density log-U(1e19, 1e20) m^-3."""
    segs = []
    for _ in range(n_fil):
        radius = rng.uniform(0.1e-3, 0.3e-3)
        dens = 10 ** rng.uniform(19, 20)
        p = np.array([rng.uniform(0.3, 0.7) * L, rng.uniform(0.3, 0.7) * L, 0.85 * L])
        pts = [p.copy()]
        for _k in range(4):
            step = np.array([rng.normal(0, 0.05), rng.normal(0, 0.05), -1.0])
            step = step / np.linalg.norm(step) * 0.1 * L
            p = p + step
            pts.append(p.copy())
        for a, b in zip(pts[:-1], pts[1:]):
            segs.append((a, b, radius, dens))
        # one side branch from the second vertex
        q = pts[2].copy()
        for _k in range(2):
            step = np.array([rng.normal(0, 0.4), rng.normal(0, 0.4), -1.0])
            step = step / np.linalg.norm(step) * 0.08 * L
            segs.append((q.copy(), q + step, radius * 0.7, dens * 0.5))
            q = q + step
    return segs


def _seg_dist(x, y, z, a, b):
    ab = b - a
    t = ((x - a[0]) * ab[0] + (y - a[1]) * ab[1] + (z - a[2]) * ab[2]) / np.dot(ab, ab)
    t = np.clip(t, 0, 1)
    return np.sqrt((x - a[0] - t * ab[0]) ** 2 + (y - a[1] - t * ab[1]) ** 2 + (z - a[2] - t * ab[2]) ** 2)


def filament_charge(segs):
    """max over segments of ``dens * exp(-(d / radius)^2)``, ``d`` the distance
    to the segment. Arguments are per-block arrays (nblk, M, M, M); blocks
    farther than 7 radii from a segment are skipped (the term is < 1e-21 of
    its peak there)."""
    def f(x, y, z):
        out = np.zeros(np.broadcast(x, y, z).shape)
        n = out.shape[0]
        cx, cy, cz = (np.asarray(v).reshape(n, -1).mean(axis=1) for v in (x, y, z))
        ext = np.asarray(x).reshape(n, -1)
        half = 0.5 * (ext.max(axis=1) - ext.min(axis=1)) * np.sqrt(3) * 1.2
        for a, b, radius, dens in segs:
            dc = _seg_dist(cx, cy, cz, a, b)
            sel = np.nonzero(dc < 7 * radius + half)[0]
            if len(sel) == 0:
                continue
            d = _seg_dist(x[sel], y[sel], z[sel], a, b)
            out[sel] = np.maximum(out[sel], dens * np.exp(-(d / radius) ** 2))
        return out
    return f


def make_cfg4(seed: int = 0, G: int = 2, scale: float = 0.85, levels: int = 5) -> Problem:
    """3D Cartesian AMR Poisson: 1 cm^3, root 8^3 blocks of 8^3 cells
    (h = 156.25 um), 5 levels, BC ((1,1),(1,1),(-1,-1)) (Neumann sides,
    Dirichlet top/bottom). Refined around three branching filaments to about
    30M cells. ``scale`` < 1 shrinks the refinement distances (for tests)."""
    rng = np.random.default_rng(seed)
    D, M = 3, 8
    rootsize = (8, 8, 8)
    L = 1e-2
    h = L / (rootsize[0] * M)
    segs = filament_segments(rng, L=L)
    dist = [1.6e-3, 1.0e-3, 0.62e-3, 0.36e-3]

    def crit(level, coords):
        hl = h / (1 << (level - 1))
        bl = M * hl
        cx = (coords - 0.5) * bl
        half = bl / 2
        dmin = np.full(len(coords), np.inf)
        for a, b, radius, dens in segs:
            d = _seg_dist(cx[:, 0], cx[:, 1], cx[:, 2], a, b)
            dmin = np.minimum(dmin, d - radius)
        return dmin < scale * dist[level - 1] + half * np.sqrt(3)

    tree = refined_tree(D, rootsize, levels, crit)
    conn = connectivity(tree)
    b = fill_field(tree, conn, M, h, G, filament_charge(segs))
    bc = boundaryconditions(((1, 1), (1, 1), (-1, -1)))
    s = h ** 2 * (ELEMENTARY_CHARGE / EPSILON_0)
    return Problem("cfg4", D, M, h, tree, conn, GEOM_CARTESIAN, bc, s, b, seed=seed)


# --------------------------------------------------------------------------
# cfg3 pieces: photoionization Helmholtz terms and random uneven trees
# --------------------------------------------------------------------------

def bourdon3_terms(pO2=150 * TORR, p=750 * TORR, pq=40 * MILLI * BAR, xi=0.075):
    """The three-term Bourdon model (src/photo.jl:94-127): returns (a, k2)."""
    A = np.array([1.986e-4, 0.0051, 0.4886])
    lam = np.array([0.0553, 0.1460, 0.8900])
    a = A * pO2 ** 2 * (pq / (p + pq)) * xi * CENTI ** -2 * TORR ** -2
    k = pO2 * lam * CENTI ** -1 * TORR ** -1
    return a, k ** 2


def photo_source(rng, n=8, rmax=5e-3, zmax=1e-2):
    """Seeded sum of ``n`` Gaussian ionization sources in (r, z)."""
    cs = [(rng.uniform(0, 0.3) * rmax, rng.uniform(0.2, 0.8) * zmax,
           rng.uniform(0.2e-3, 0.8e-3), 10 ** rng.uniform(24, 26)) for _ in range(n)]

    def f(r, z):
        out = np.zeros(np.broadcast(r, z).shape)
        for r0, z0, w, A in cs:
            out = out + A * np.exp(-(((r - r0) ** 2 + (z - z0) ** 2) / w ** 2))
        return out
    return f


def make_helmholtz_cfg1(seed: int = 0, G: int = 2) -> List[Problem]:
    """The cfg1 tree with the three Bourdon Helmholtz terms: k^2 from
    ``bourdon3``, s = h^2, BC ((1,1),(-1,-1)) (src/photo.jl:108-110), source
    = seeded sum of 8 Gaussians. One Problem per term, sharing tree and b."""
    base = make_cfg1(seed, G)
    rng = np.random.default_rng(1000 + seed)
    b = fill_field(base.tree, base.conn, base.M, base.h, G, photo_source(rng))
    _, k2s = bourdon3_terms()
    bc = boundaryconditions(((1, 1), (-1, -1)))
    out = []
    for i, k2 in enumerate(k2s):
        out.append(Problem(f"cfg3-term{i}", 2, base.M, base.h, base.tree, base.conn,
                           GEOM_CYL1, bc, base.h ** 2, b, k2=float(k2), helm=True, seed=seed))
    return out


def make_random_tree_problem(seed: int, G: int = 2, min_cells: float = 1e4,
                             max_cells: float = 1e6, helm_k2: Optional[float] = None) -> Problem:
    """An uneven 2D (r, z) tree: root between 1x2 and 4x8 blocks of 16x16,
    random depth 2..6, refined around random sources until its size lands
    log-uniformly in [min_cells, max_cells]."""
    rng = np.random.default_rng(seed)
    D, M = 2, 16
    rx = int(rng.integers(1, 5))
    rootsize = (rx, 2 * rx)
    depth = int(rng.integers(2, 7))
    target = 10 ** rng.uniform(np.log10(min_cells), np.log10(max_cells))
    L = 5e-3
    h = L / (rootsize[0] * M)
    srcs = [(rng.uniform(0, 0.5) * L, rng.uniform(0.2, 0.8) * 2 * L) for _ in range(3)]

    def build(radius):
        def crit(level, coords):
            hl = h / (1 << (level - 1))
            bl = M * hl
            cx = (coords - 0.5) * bl
            ok = np.zeros(len(coords), bool)
            for r0, z0 in srcs:
                d = np.sqrt((cx[:, 0] - r0) ** 2 + (cx[:, 1] - z0) ** 2)
                ok |= d < radius * 0.6 ** (level - 1) + bl
            return ok
        return refined_tree(D, rootsize, depth, crit)

    lo, hi = 0.0, 2 * L
    tree = build(hi)
    for _ in range(18):
        mid = 0.5 * (lo + hi)
        t = build(mid)
        n = t.nblocks() * M * M
        if n > target:
            hi = mid
        else:
            lo, tree = mid, t
    conn = connectivity(tree)
    f = photo_source(rng, n=8, rmax=L, zmax=2 * L)
    b = fill_field(tree, conn, M, h, G, f)
    if helm_k2 is None:
        bc = boundaryconditions(((1, -1), (-1, -1)))
        s = h ** 2 * (ELEMENTARY_CHARGE / EPSILON_0)
        return Problem(f"tree{seed}", D, M, h, tree, conn, GEOM_CYL1, bc, s, b, seed=seed)
    bc = boundaryconditions(((1, 1), (-1, -1)))
    return Problem(f"tree{seed}", D, M, h, tree, conn, GEOM_CYL1, bc, h ** 2, b,
                   k2=helm_k2, helm=True, seed=seed)


def uniform_problem(D: int, rootsize: Sequence[int], M: int, levels: int, geom: int,
                    bc_desc, helm: bool = False, k2: float = 0.0, seed: int = 0,
                    G: int = 2, refine: Optional[Callable] = None) -> Problem:
    """Small test problems: a (possibly refined) tree with a random smooth
    right-hand side; used by the parity and known-answer tests."""
    rng = np.random.default_rng(seed)
    crit = refine if refine is not None else (lambda l, c: np.zeros(len(c), bool))
    tree = refined_tree(D, rootsize, levels, crit)
    conn = connectivity(tree)
    h = 1.0 / (rootsize[0] * M)
    cen = rng.uniform(0.2, 0.8, size=(3, D))
    amp = rng.uniform(-1, 1, size=3)

    def f(*x):
        out = np.zeros(np.broadcast(*x).shape)
        for c, a in zip(cen, amp):
            r2 = sum((x[d] - c[d] * rootsize[d] / rootsize[0]) ** 2 for d in range(D))
            out = out + a * np.exp(-r2 / 0.02)
        return out
    b = fill_field(tree, conn, M, h, G, f)
    bc = boundaryconditions(bc_desc)
    s = h ** 2
    return Problem(f"uniform{D}d", D, M, h, tree, conn, geom, bc, s, b,
                   k2=k2 if helm else 0.0, helm=helm, seed=seed)
