"""Host block fields in the reference's own memory layout.

``ScalarBlockField{D,M,G,T}`` with contiguous storage (src/blockfield.jl:33-59)
is an ``ElasticArray{T}(S, ..., S, nblk)``: column-major, first index fastest,
block id (1-based) last, ``S = M + 2G``. Here that is a Fortran-ordered numpy
array of shape ``(S,)*D + (nblk,)`` so that ``u[i1, i2, blk-1]`` addresses the
same byte as Julia's ``u.u[i1+1, i2+1, blk]``. This is exactly the buffer a
Julia caller passes through the C ABI.
"""

from __future__ import annotations

import numpy as np


class ScalarBlockField:
    def __init__(self, D: int, M: int, G: int, nblocks: int, dtype=np.float64):
        if M % 2:
            raise ValueError("M must be even (src/blockfield.jl:52)")
        self.D, self.M, self.G = D, M, G
        self.S = M + 2 * G
        self.u = np.zeros((self.S,) * D + (nblocks,), dtype=dtype, order="F")

    @property
    def nblocks(self) -> int:
        return self.u.shape[-1]

    def block(self, blk: int) -> np.ndarray:
        """View of block ``blk`` (1-based), ghosts included."""
        return self.u[..., blk - 1]

    def valid(self, blk: int) -> np.ndarray:
        """View of the valid cells of block ``blk`` (src/blockfield.jl:399-401)."""
        G, M = self.G, self.M
        return self.u[(slice(G, G + M),) * self.D + (blk - 1,)]

    def copy(self) -> "ScalarBlockField":
        f = ScalarBlockField(self.D, self.M, self.G, self.nblocks, self.u.dtype)
        f.u[...] = self.u
        return f

    def like(self) -> "ScalarBlockField":
        return ScalarBlockField(self.D, self.M, self.G, self.nblocks, self.u.dtype)


def cell_centers(coord, level: int, M: int, h: float):
    """src/mesh.jl:111-116: per-dimension cell-centre coordinates of a block."""
    hl = h / (1 << (level - 1))
    return [hl * (M * (c - 1) + np.arange(1, M + 1)) - hl / 2 for c in coord]


def restrict_full_host(f: ScalarBlockField, tree, conn) -> None:
    """Host-side ``restrict_full!`` (src/restrict.jl:5-47) used to build
    right-hand sides: every covered coarse cell becomes the mean of its 2^D
    children, finest level first. Summation in the reference's loop order
    (first index fastest)."""
    D, M, G = f.D, f.M, f.G
    h = M // 2
    vs = (slice(G, G + M),) * D
    for li in range(len(conn.child) - 1, -1, -1):
        rec = conn.child[li]
        if len(rec) == 0:
            continue
        fine = rec[:, 0] - 1
        coarse = rec[:, 1] - 1
        sb = rec[:, 2:2 + D]
        src = f.u[vs + (fine,)]
        acc = np.zeros((h,) * D + (len(rec),), dtype=f.u.dtype)
        for offs in np.ndindex(*(2,) * D):
            # the reference's loop has i_1 innermost -> reversed ndindex
            o = tuple(reversed(offs))
            acc = acc + src[tuple(slice(o[d], M, 2) for d in range(D)) + (slice(None),)]
        acc = acc / (1 << D)
        for pat in np.ndindex(*(2,) * D):
            sel = np.all(sb == (np.array(pat) + 1), axis=1)
            if not sel.any():
                continue
            sl = tuple(slice(G + pat[d] * h, G + (pat[d] + 1) * h) for d in range(D))
            f.u[sl + (coarse[sel],)] = acc[..., sel]
