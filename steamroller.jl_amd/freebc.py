"""Free boundary conditions of the 2D cylindrical Poisson problem
(src/freebc.jl, following Malagon-Romero 2018).

The radial boundary r = R is made open: the potential is solved once with the
homogeneous Dirichlet condition there (``uh``), the radial derivative at the
boundary is expanded in the z sine series, each mode is mapped to its boundary
value through the modified-Bessel kernel f, and the boundary values enter the
right-hand side (``setouter!``), so that the following solve with the same
homogeneous condition gives the free-boundary potential. ``fill_ghost_free!``
adds the boundary values to the ghosts used by the electric field.

The kernel f is computed here, on the host, as the reference's constructor does
(src/freebc.jl:32-54, with SpecialFunctions' scaled Bessel functions; scipy's
``ive``/``kve`` are the same functions); the transforms and the field updates
run in the library (``srmg_freebc_*``).
"""

from __future__ import annotations

import ctypes as C
import math
from typing import Sequence

import numpy as np

from . import solver
from .solver import _check, _p, lib


def boundary_kernel(n: int, H: float, R: float) -> np.ndarray:
    """f_m = (2/H) / (k (I1(kR)/I0(kR) + K1(kR)/K0(kR))), k = pi m / H, m = 1..n
    (src/freebc.jl:38-42); the exponential scalings of I and K cancel in the
    ratios."""
    from scipy.special import ive, kve
    k = (math.pi / H) * np.arange(1, n + 1, dtype=np.float64)
    x = k * R
    return (2.0 / H) / (k * (ive(1, x) / ive(0, x) + kve(1, x) / kve(0, x)))


class FreeBC:
    """``FreeBC(geom::CylindricalGeometry{dim}, u, level, rootsize, h)``
    (src/freebc.jl:32-54) on a context: ``level`` is the level at which the
    boundary is resolved (the reference passes ``refine_maxlevel``), ``rootsize``
    the level-1 block domain, ``h`` the level-1 cell size. ``uh`` is the
    homogeneous potential (a device field of the same context) that
    :meth:`setfreebc` keeps between calls, as the reference keeps its own."""

    def __init__(self, ctx: solver.Context, geometry: solver.CylindricalGeometry, level: int,
                 rootsize: Sequence[int], h: float):
        if not isinstance(geometry, solver.CylindricalGeometry) or ctx.D != 2:
            raise solver.SrmgError("free boundary conditions need a 2D cylindrical geometry")
        self.ctx = ctx
        self.geometry = geometry
        self.dim = int(geometry.N)
        other = self.dim % 2 + 1
        self.level = int(level)
        self.rootsize = tuple(int(r) for r in rootsize)
        self.h = float(h)
        M = ctx.M
        self.n = self.rootsize[other - 1] * 2 ** (self.level - 1) * M
        self.H = self.rootsize[other - 1] * self.h * M
        self.R = self.rootsize[self.dim - 1] * self.h * M
        self.f = boundary_kernel(self.n, self.H, self.R)
        self.uh = ctx.field(solver.F64)
        self.uh.zero()
        root = np.asarray(self.rootsize, np.int64)
        hdl = C.c_void_p()
        _check(lib().srmg_freebc_create(ctx.h, C.byref(hdl), self.dim, self.level, _p(root), self.h,
                                        _p(self.f), int(self.n)))
        self.hdl = hdl

    def __del__(self):
        h = getattr(self, "hdl", None)
        if h is not None and h.value:
            try:
                lib().srmg_freebc_destroy(h)
            except Exception:
                pass
            self.hdl = None

    def freebc(self, q: solver.DeviceField, u: solver.DeviceField, s: float):
        """freebc!(q, u, h, s, tree, conn, conf) (src/freebc.jl:87-110): q gains
        the boundary sources of the free boundary computed from the homogeneous
        potential u."""
        _check(lib().srmg_freebc_apply(self.ctx.h, self.hdl, q.id, u.id, float(s)))

    def setfreebc(self, q1: solver.DeviceField, r: solver.DeviceField, u1: solver.DeviceField, s: float, pbc,
                  lpl, *, poisson_iter: int = 1, nup: int = 2, ndown: int = 2, ntop: int = 4,
                  poisson_fmg: bool = False):
        """setfreebc!(fr, pbc, strfields, strconf, tree, conn) (src/freebc.jl:59-82):
        ``poisson_iter`` cycles on the homogeneous potential (warm start), then
        :meth:`freebc` into q1. Returns pbc, as the reference does."""
        for _ in range(int(poisson_iter)):
            solver.fill_ghost_all(self.uh, pbc)
            if poisson_fmg:
                solver.fmg(self.uh, q1, r, u1, s, geometry=self.geometry, bc=pbc, lpldisc=lpl, nup=nup,
                           ndown=ndown, ntop=ntop)
            else:
                solver.vcycle(self.uh, q1, r, u1, s, geometry=self.geometry, bc=pbc, lpldisc=lpl, nup=nup,
                              ndown=ndown, ntop=ntop)
        self.freebc(q1, self.uh, s)
        return pbc

    def boundary(self) -> np.ndarray:
        """The boundary values b of the last :meth:`freebc` (n values along the
        other dimension at the resolution of ``level``)."""
        b = np.empty(self.n)
        _check(lib().srmg_freebc_boundary(self.hdl, _p(b), int(self.n)))
        return b

    def electric_field(self, u: solver.DeviceField, eabs: solver.DeviceField, h: float, e0, bc, G: int = 2,
                       with_faces: bool = True):
        """:func:`solver.electric_field` after the ghost fill of fill_ghost_free!
        (src/freebc.jl:218-238, caller src/streamer.jl:340-349)."""
        c = u.ctx
        e = np.ascontiguousarray(np.asarray(e0, np.float64))
        b = solver._bc_arr(bc)
        out = None
        if with_faces:
            S = c.M + 2 * G + 1
            out = np.zeros((S,) * c.D + (c.D, c.nblocks), order="F")
        _check(lib().srmg_freebc_electric_field(c.h, self.hdl, u.id, eabs.id, float(h), _p(e), _p(b),
                                                None if out is None else out.ctypes.data_as(C.c_void_p), int(G)))
        return out
