"""ctypes binding of libsrmg.so and a Python mirror of the reference's solver API.

The reference's operator plug-in is Julia dispatch on a discretization and a
geometry type (src/laplacian.jl:9-14, src/helmholtz.jl:9-11, src/geometry.jl:14-20);
the same values are carried here by small descriptor classes and passed through
the C ABI as an ``srmg_op``. Fields are device-resident (:class:`DeviceField`);
the free functions ``vcycle``/``fmg``/``resnorm``/``fill_ghost`` keep the
reference's argument order and meaning (src/multigrid.jl:17-106, 391-417;
src/connectivity.jl:155-159).

There is no CPU fallback: if the HIP library is missing or no GPU is present
the calls raise.

When torch shares the process (the learned first iterate's training data,
``grid_solve`` and ``UNet`` on torch tensors), import torch before the first
library call: torch ships its own HIP runtime, and it finds no GPU when
another HIP runtime was initialised first.
"""

from __future__ import annotations

import ctypes as C
import os
import time
from dataclasses import dataclass
from typing import Optional

import numpy as np

from .tree import flatten_pairs, flatten_records

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsrmg.so")

F64, F32 = 0, 1
KERNEL_GS_PASS, KERNEL_RESIDUAL, KERNEL_GS_SWEEP, KERNEL_STREAM, KERNEL_RESIDUAL_GATHER, KERNEL_STREAM16 = 0, 1, 2, 3, 4, 5
SWEEP_SHAPES = 12  # tile shapes of the fused smoother (Context.set_sweep_config)
COARSE_GS, COARSE_EXACT = 0, 1
REGRID_INTERP, REGRID_RESTRICT = 1, 2  # srmg_regrid modes
GUESS_NONE, GUESS_LEARNED, GUESS_COARSE, GUESS_ZERO = 0, 1, 2, 3
_DT = {F64: np.float64, F32: np.float32}


class SrmgError(RuntimeError):
    pass


# ---- operator descriptors (mirror the reference's types) ----

@dataclass(frozen=True)
class CartesianGeometry:
    code: int = 0


@dataclass(frozen=True)
class CylindricalGeometry:
    """``CylindricalGeometry{N}``: dimension N (1-based) is the radius."""
    N: int = 1

    @property
    def code(self) -> int:
        return self.N


@dataclass(frozen=True)
class LaplacianDiscretization:
    D: int
    P: int = 2
    k2: float = 0.0

    @property
    def code(self) -> int:
        return 2 if self.P == 4 else 0  # SRMG_LAPLACIAN4 / SRMG_LAPLACIAN


@dataclass(frozen=True)
class HelmholtzDiscretization:
    D: int
    P: int
    k2: float
    code: int = 1


def geometry_from_code(code: int):
    return CartesianGeometry() if code == 0 else CylindricalGeometry(code)


class _Op(C.Structure):
    _fields_ = [("geometry", C.c_int32), ("disc", C.c_int32), ("k2", C.c_double),
                ("bc", C.c_int8 * 27), ("coarse", C.c_int32)]


class _SolveOpts(C.Structure):
    _fields_ = [("ndown", C.c_int32), ("ntop", C.c_int32), ("nup", C.c_int32),
                ("omega", C.c_double), ("tol", C.c_double), ("max_cycles", C.c_int32),
                ("use_graph", C.c_int32), ("inner_cycles", C.c_int32), ("guess", C.c_int32),
                ("guess_smooth", C.c_int32), ("first_fmg", C.c_int32), ("inner_first", C.c_int32)]


def make_op(geometry, bc, disc, coarse: int = COARSE_GS, any_order: bool = False) -> _Op:
    if getattr(disc, "P", 2) != 2 and not any_order:
        raise SrmgError("only order-2 discretizations are implemented on the device")
    op = _Op()
    op.geometry = int(geometry.code)
    op.disc = int(disc.code)
    op.k2 = float(disc.k2)
    b = np.asarray(bc, np.int8).ravel()
    for i in range(len(b)):
        op.bc[i] = int(b[i])
    op.coarse = int(coarse)
    return op


# ---- library loading ----

_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SrmgError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    P, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "srmg_create": [C.POINTER(P), i32, i32],
        "srmg_destroy": [P],
        "srmg_set_tree": [P, i32, i64] + [P] * 10,
        "srmg_regrid": [P, i32, i64] + [P] * 10 + [P, P, i32, P],
        "srmg_regrid_forest": [P, i32, P, P] + [P] * 10 + [P, P, i32, P],
        "srmg_connectivity": [i32, i32, P, P, P, i32] + [P, P, i64] * 4,
        "srmg_restrict_full": [P, i32],
        "srmg_fill_ghost4": [P, i32, P],
        "srmg_rhs": [P, i32, i32, P],
        "srmg_field_create": [P, i32, C.POINTER(i32)],
        "srmg_field_destroy": [P, i32],
        "srmg_upload": [P, i32, P, i32, i32],
        "srmg_download": [P, i32, P, i32, i32],
        "srmg_download_bc": [P, i32, P, i32, i32, P],
        "srmg_zero": [P, i32],
        "srmg_fill_ghost": [P, i32, i32, P],
        "srmg_fill_ghost_all": [P, i32, P],
        "srmg_vcycle": [P, i32, i32, i32, i32, dbl, P, i32, i32, i32, i32, dbl],
        "srmg_fmg": [P, i32, i32, i32, i32, dbl, P, i32, i32, i32, dbl],
        "srmg_resnorm": [P, i32, i32, i32, dbl, P, C.POINTER(dbl)],
        "srmg_synchronize": [P],
        "srmg_solve": [P, i32, i32, i32, i32, dbl, P, P, C.POINTER(i32), P],
        "srmg_solve_mixed": [P, i32, i32, i32, i32, i32, i32, dbl, P, P, C.POINTER(i32), P],
        "srmg_axis_decomp": [i32, i32, dbl, dbl, P, P, P],
        "srmg_time_kernel": [P, i32, i32, i32, i32, i32, dbl, P, i32, C.POINTER(dbl)],
        "srmg_set_forest": [P, i32, P, P] + [P] * 10,
        "srmg_tree_count": [P, C.POINTER(i32)],
        "srmg_vcycle_batched": [P, i32, i32, i32, i32, P, P, i32, i32, i32, i32, dbl],
        "srmg_resnorm_batched": [P, i32, i32, i32, P, P, P],
        "srmg_solve_batched": [P, i32, i32, i32, i32, P, P, P, C.POINTER(i32), P],
        "srmg_solve_mixed_batched": [P, i32, i32, i32, i32, i32, i32, P, P, P, C.POINTER(i32), P],
        "srmg_level_info": [P, i32, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)],
        "srmg_autotune": [P, i32, i32, dbl, P, i32, P],
        "srmg_set_sweep_config": [P, i32, i32, i32],
        "srmg_level_grouped": [P, i32, C.POINTER(i32)],
        "srmg_set_group_tiles": [P, i32],
        "srmg_set_axis_gemm": [P, i32],
        "srmg_set_group_order": [P, i32],
        "srmg_electric_field": [P, i32, i32, dbl, P, P, P, i32],
        "srmg_freebc_create": [P, C.POINTER(P), i32, i32, P, dbl, P, i64],
        "srmg_freebc_destroy": [P],
        "srmg_freebc_apply": [P, P, i32, i32, dbl],
        "srmg_freebc_boundary": [P, P, i64],
        "srmg_freebc_electric_field": [P, P, i32, i32, dbl, P, P, P, i32],
        "srmg_set_guess": [P, P],
        "srmg_guess_fallbacks": [P, C.POINTER(i32)],
        "srmg_grid_solve": [P, i32, P, dbl, P, P, P, i32],
        "srmg_grid_solve_noalloc": [P, i32, P, dbl, P, P, P, i32],
        "srmg_unet_create": [C.POINTER(P), i32, P, i32, P, C.c_uint64],
        "srmg_unet_destroy": [P],
        "srmg_unet_param_count": [P, C.POINTER(i64)],
        "srmg_unet_get_params": [P, P],
        "srmg_unet_set_params": [P, P],
        "srmg_unet_set_output_scale": [P, dbl],
        "srmg_unet_get_output_scale": [P, C.POINTER(dbl)],
        "srmg_unet_set_problem": [P, P],
        "srmg_unet_forward": [P, P, P, i32],
        "srmg_unet_loss_grad": [P, P, P, i32, P, P],
        "srmg_unet_train_step": [P, P, P, i32, C.c_float, P],
        "srmg_unet_synchronize": [P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.srmg_last_error.restype = C.c_char_p
    L.srmg_last_error.argtypes = []
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        msg = lib().srmg_last_error().decode(errors="replace")
        raise SrmgError(f"srmg error {rc}: {msg}")


def _p(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def axis_decomp(n: int, radial: bool, alow: float, ahigh: float):
    """Host-only: the level-1 1D operator as P diag(lam) Q (no device work)."""
    lam = np.zeros(n)
    P = np.zeros((n, n))
    Q = np.zeros((n, n))
    _check(lib().srmg_axis_decomp(n, int(radial), alow, ahigh, _p(lam), _p(P), _p(Q)))
    return lam, P, Q


# ---- connectivity! on the device ----

STAR_STENCIL, BOX_STENCIL = 0, 1


def connectivity_device(tree, stencil: int = STAR_STENCIL):
    """connectivity!(conn, tree, stencil) (src/connectivity.jl:106-149) computed
    by ``srmg_connectivity`` on the GPU; returns the same
    :class:`srmg.tree.Connectivity` as :func:`srmg.tree.connectivity`."""
    from .tree import Connectivity
    D, L = tree.D, len(tree)
    pairs, poff = flatten_pairs(tree)
    n = int(poff[-1])
    ns = 2 * D if stencil == STAR_STENCIL else 3 ** D - 1
    cap = max(n * ns, 1)
    nchild = max(n - int(poff[1]), 1) if L > 0 else 1
    root = np.array(tree.rootsize, np.int64)
    nb, bd, rb = (np.zeros((cap, w), np.int64) for w in (2 + D, 2 + 2 * D, 2 + 2 * D))
    ch = np.zeros((nchild, 2 + D), np.int64)
    offs = [np.zeros(L + 1, np.int64) for _ in range(4)]
    _check(lib().srmg_connectivity(D, L, _p(root), _p(pairs), _p(poff), int(stencil),
                                   _p(nb), _p(offs[0]), cap, _p(bd), _p(offs[1]), cap,
                                   _p(rb), _p(offs[2]), cap, _p(ch), _p(offs[3]), nchild))
    conn = Connectivity(D, L)
    for lists, arr, o in ((conn.neighbor, nb, offs[0]), (conn.boundary, bd, offs[1]),
                          (conn.refboundary, rb, offs[2]), (conn.child, ch, offs[3])):
        for l in range(L):
            lists[l] = arr[o[l]:o[l + 1]].copy()
    return conn


# ---- device objects ----

class Context:
    """One device context bound to a tree (``srmg_create`` + ``srmg_set_tree``)."""

    def __init__(self, tree, conn, M: int, forest=None, nblocks: int | None = None):
        self.D, self.M = tree.D, M
        self.tree, self.conn = tree, conn
        h = C.c_void_p()
        _check(lib().srmg_create(C.byref(h), tree.D, M))
        self.h = h
        if forest is None:
            self.ntrees = 1
            self.regrid(tree, conn, nblocks=nblocks)
        else:
            self.set_forest(forest)

    @classmethod
    def forest(cls, trees_conns, M: int) -> "Context":
        """One context for several independent trees (``srmg_set_forest``);
        host fields are the trees' host arrays concatenated in order
        (see :func:`concat_fields`)."""
        t0, c0 = trees_conns[0]
        return cls(t0, c0, M, forest=list(trees_conns))

    def set_forest(self, trees_conns, fields=(), bc=None, nblocks=None):
        """Install a forest (``srmg_set_forest``); with ``fields`` the trees
        replace the context's trees one for one and those fields are carried
        across (``srmg_regrid_forest``, modes as in :meth:`regrid`).
        ``nblocks``: host block ids per tree (default: the trees' counts)."""
        D = trees_conns[0][0].D
        keep, per = [], [[] for _ in range(10)]
        nl, nbk = [], []
        for ti, (tree, conn) in enumerate(trees_conns):
            pairs, poff = flatten_pairs(tree)
            arrs = [pairs, poff]
            for lists, w in ((conn.neighbor, 2 + D), (conn.boundary, 2 + 2 * D),
                             (conn.refboundary, 2 + 2 * D), (conn.child, 2 + D)):
                a, o = flatten_records(lists, w)
                if len(a) == 0:
                    a = np.zeros((1, w), np.int64)
                arrs += [a, o]
            keep.append(arrs)
            for k in range(10):
                per[k].append(arrs[k].ctypes.data)
            nl.append(len(tree))
            nbk.append(tree.nblocks() if nblocks is None else int(nblocks[ti]))
        ptrs = [(C.c_void_p * len(v))(*v) for v in per]
        nl_a = np.array(nl, np.int32)
        nb_a = np.array(nbk, np.int64)
        self._tables = (keep, ptrs, nl_a, nb_a)
        fields = list(fields)
        if fields:
            ids = np.array([f.id for f, _ in fields], np.int32)
            modes = np.array([int(m) for _, m in fields], np.int32)
            b = None if bc is None else _bc_arr(bc)
            _check(lib().srmg_regrid_forest(self.h, len(trees_conns), _p(nl_a), _p(nb_a), *ptrs, _p(ids),
                                            _p(modes), len(fields), None if b is None else _p(b)))
        else:
            _check(lib().srmg_set_forest(self.h, len(trees_conns), _p(nl_a), _p(nb_a), *ptrs))
        self.ntrees = len(trees_conns)
        self.trees = [t for t, _ in trees_conns]
        self.nlevels = int(nl_a.max())
        self.nblocks = int(nb_a.sum())

    def regrid(self, tree, conn, fields=(), bc=None, nblocks: int | None = None):
        """Install a new tree (``srmg_set_tree``). ``nblocks`` is the number of
        host block ids (the host fields' length; default: the tree's block
        count). Without ``fields`` every device field is reallocated at zero.
        ``fields`` = [(DeviceField, mode), ...] carries those fields across on
        the device (``srmg_regrid``: applydelta!'s interp!/restrict!, mode a
        combination of REGRID_INTERP and REGRID_RESTRICT); ``bc`` is needed
        when a field is interpolated."""
        D = tree.D
        nblocks = tree.nblocks() if nblocks is None else int(nblocks)
        pairs, poff = flatten_pairs(tree)
        arrs = [pairs, poff]
        for lists, w in ((conn.neighbor, 2 + D), (conn.boundary, 2 + 2 * D),
                         (conn.refboundary, 2 + 2 * D), (conn.child, 2 + D)):
            a, o = flatten_records(lists, w)
            if len(a) == 0:
                a = np.zeros((1, w), np.int64)
            arrs += [a, o]
        self._tables = arrs
        fields = list(fields)
        t0 = time.perf_counter()
        if fields:
            ids = np.array([f.id for f, _ in fields], np.int32)
            modes = np.array([int(m) for _, m in fields], np.int32)
            b = None if bc is None else _bc_arr(bc)
            _check(lib().srmg_regrid(self.h, len(tree), nblocks, *[_p(a) for a in arrs], _p(ids), _p(modes),
                                     len(fields), None if b is None else _p(b)))
        else:
            _check(lib().srmg_set_tree(self.h, len(tree), nblocks, *[_p(a) for a in arrs]))
        # wall time of the library call alone (both return after the device work)
        self.last_regrid_ms = 1e3 * (time.perf_counter() - t0)
        self.tree, self.conn = tree, conn
        self.ntrees = 1
        self.nlevels = len(tree)
        self.nblocks = nblocks

    def level_info(self, level: int):
        """(blocks, leaf blocks, coarse-fine ghost slabs) of a 1-based level."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        _check(lib().srmg_level_info(self.h, level, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def time_kernel(self, kind: int, level: int, u, b, r, s, geometry, bc, disc, reps=20) -> float:
        """Average ms per launch of a level kernel (HIP events on the context stream)."""
        op = make_op(geometry, bc, disc)
        ms = C.c_double()
        _check(lib().srmg_time_kernel(self.h, kind, level, u.id, b.id, r.id, float(s), C.byref(op),
                                      int(reps), C.byref(ms)))
        return ms.value

    def autotune(self, u, b, s, geometry, bc, disc, reps=5):
        """Pick the fastest fused-sweep tile shape per level (times the sweeps on u;
        on a forest every tree takes s and disc for the trial sweeps)."""
        op = make_op(geometry, bc, disc)
        out = np.zeros(self.nlevels, np.int32)
        _check(lib().srmg_autotune(self.h, u.id, b.id, float(s), C.byref(op), int(reps), _p(out)))
        return out

    def level_grouped(self, level: int) -> bool:
        g = C.c_int()
        _check(lib().srmg_level_grouped(self.h, int(level), C.byref(g)))
        return bool(g.value)

    def set_group_tiles(self, enable: bool):
        _check(lib().srmg_set_group_tiles(self.h, int(bool(enable))))

    def set_axis_gemm(self, variant: int):
        """Kernels of the level-1 solve: 2 (default) three launches for 3D
        grids with extents that are multiples of 8 and at most 64 (else as 1),
        1 one launch per axis transform with one wave per work group, 0 the
        LDS-tiled axis kernels (all give the same bits)."""
        _check(lib().srmg_set_axis_gemm(self.h, int(variant)))

    def set_group_order(self, order: int):
        """Order in which the column sweeps of grouped 3D levels visit the
        sibling groups: 0 (default) the device block order; 1 slabs two
        groups thick in z, row by row (x and y neighbours close in time, less
        HBM traffic); same bits either way."""
        _check(lib().srmg_set_group_order(self.h, int(order)))

    def set_sweep_config(self, dtype: int, level: int, cfg: int):
        _check(lib().srmg_set_sweep_config(self.h, dtype, level, cfg))

    def field(self, dtype: int = F64) -> "DeviceField":
        i = C.c_int()
        _check(lib().srmg_field_create(self.h, dtype, C.byref(i)))
        return DeviceField(self, i.value, dtype)

    def synchronize(self):
        _check(lib().srmg_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().srmg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceField:
    def __init__(self, ctx: Context, fid: int, dtype: int):
        self.ctx, self.id, self.dtype = ctx, fid, dtype

    def upload(self, host) -> "DeviceField":
        a = host.u if hasattr(host, "u") else host
        if not a.flags.f_contiguous:
            raise SrmgError("host field must be Fortran-ordered (the reference's layout)")
        G = (a.shape[0] - self.ctx.M) // 2
        hd = F64 if a.dtype == np.float64 else F32
        if a.dtype not in (np.float64, np.float32):
            raise SrmgError("host field must be float64 or float32")
        _check(lib().srmg_upload(self.ctx.h, self.id, _p(a), hd, G))
        return self

    def download(self, host, bc=None):
        a = host.u if hasattr(host, "u") else host
        if not a.flags.f_contiguous:
            raise SrmgError("host field must be Fortran-ordered (the reference's layout)")
        G = (a.shape[0] - self.ctx.M) // 2
        hd = F64 if a.dtype == np.float64 else F32
        if bc is None:
            _check(lib().srmg_download(self.ctx.h, self.id, _p(a), hd, G))
        else:
            bcv = np.ascontiguousarray(np.asarray(bc, np.int8).ravel())
            _check(lib().srmg_download_bc(self.ctx.h, self.id, _p(a), hd, G, _p(bcv)))
        return host

    def zero(self):
        _check(lib().srmg_zero(self.ctx.h, self.id))
        return self

    def close(self):
        """Release the device buffers (the id may be reused by a later field)."""
        if self.id is not None and getattr(self.ctx, "h", None):
            _check(lib().srmg_field_destroy(self.ctx.h, self.id))
        self.id = None


# ---- the reference's entry points ----

def _bc_arr(bc):
    return np.ascontiguousarray(np.asarray(bc, np.int8).ravel())


def fill_ghost(u: DeviceField, l: int, conn=None, bc=None):
    """fill_ghost!(u, l, conn, bc) (src/connectivity.jl:155-159)."""
    b = _bc_arr(bc)
    _check(lib().srmg_fill_ghost(u.ctx.h, u.id, l, _p(b)))


def restrict_full(u: DeviceField):
    """restrict_full!(u, conn) (src/restrict.jl:43-47) on the device."""
    _check(lib().srmg_restrict_full(u.ctx.h, u.id))


def rhs(b1: DeviceField, b: DeviceField, geometry, bc, lpldisc):
    """derivs!'s right-hand side loop (src/streamer.jl:314-318): for every
    level, fill_ghost!(b, l) and rhs_level!(b1, b, l) (src/multigrid.jl:378-388);
    b1 = b for order 2, the compact R b for LaplacianDiscretization(D, 4)."""
    op = make_op(geometry, bc, lpldisc, any_order=True)
    _check(lib().srmg_rhs(b.ctx.h, b1.id, b.id, C.byref(op)))


def extrapolate_const(D: int) -> np.ndarray:
    """``ExtrapolateConst()`` (src/boundary.jl:18-21, 49) as a boundary table:
    its getbc returns the mirror cell's value, i.e. every entry +1. The
    reference fills |E|'s ghosts with it after ``electric_field!``
    (src/streamer.jl:353-355): ``fill_ghost_all(eabs, extrapolate_const(D))``,
    then ``eabs.download(host, bc=extrapolate_const(D))``."""
    return np.ones((3,) * D, np.int8)


def fill_ghost_all(u: DeviceField, bc, lpldisc=None):
    """fill_ghost!(u, l, conn, bc) for every level (src/streamer.jl:323-325);
    with LaplacianDiscretization(D, 4) the order-4 ghosts (corners and edges
    included, ``srmg_fill_ghost4``)."""
    b = _bc_arr(bc)
    if lpldisc is not None and getattr(lpldisc, "P", 2) == 4:
        _check(lib().srmg_fill_ghost4(u.ctx.h, u.id, _p(b)))
        return
    _check(lib().srmg_fill_ghost_all(u.ctx.h, u.id, _p(b)))


def vcycle(u, b, r, u1, s, tree=None, conn=None, geometry=None, bc=None, lpldisc=None, *,
           lmax=0, ndown=3, ntop=5, nup=3, omega=1.0, coarse=COARSE_GS):
    """vcycle! (src/multigrid.jl:60-106) on device-resident fields; with
    LaplacianDiscretization(D, 4) the order-4 form (Cartesian or cylindrical,
    FP64, complete levels: FourColors smoothing, src/multigrid.jl:268-295)."""
    op = make_op(geometry, bc, lpldisc, coarse, any_order=True)
    _check(lib().srmg_vcycle(u.ctx.h, u.id, b.id, r.id, u1.id, float(s), C.byref(op),
                             int(lmax), int(ndown), int(ntop), int(nup), float(omega)))


def fmg(u, b, r, u1, s, tree=None, conn=None, geometry=None, bc=None, lpldisc=None, *,
        lmax=0, ndown=3, ntop=5, nup=3, omega=1.25, coarse=COARSE_GS):
    """fmg! (src/multigrid.jl:17-51) on device-resident fields (order 4 as
    :func:`vcycle`)."""
    op = make_op(geometry, bc, lpldisc, coarse, any_order=True)
    _check(lib().srmg_fmg(u.ctx.h, u.id, b.id, r.id, u1.id, float(s), C.byref(op),
                          int(ndown), int(ntop), int(nup), float(omega)))


def poisson_step(u, q1, r, u1, s, geometry, pbc, lpl, *, poisson_iter=1, poisson_fmg=False, nup=2, ndown=2,
                 ntop=4, freebc=None, coarse=COARSE_GS, q=None):
    """The Poisson block of derivs! (src/streamer.jl:309-333). With the net
    charge ``q`` given: restrict_full!(q), then per level fill_ghost!(q) and
    q1 = rhs(q) (313-318); without it q1 already holds the right-hand side.
    Then setfreebc! when a free boundary is given (``srmg.freebc.FreeBC``), and
    ``poisson_iter`` times the ghost pre-fill of u on every level and one fmg!
    (``poisson_fmg``) or vcycle!, warm-started from the previous u.
    (nup, ndown, ntop) is the reference's poisson_level_iter."""
    if q is not None:
        restrict_full(q)
        rhs(q1, q, geometry, pbc, lpl)
    if freebc is not None:
        freebc.setfreebc(q1, r, u1, s, pbc, lpl, poisson_iter=poisson_iter, nup=nup, ndown=ndown, ntop=ntop,
                         poisson_fmg=poisson_fmg)
    for _ in range(int(poisson_iter)):
        fill_ghost_all(u, pbc, lpl)
        if poisson_fmg:
            fmg(u, q1, r, u1, s, geometry=geometry, bc=pbc, lpldisc=lpl, nup=nup, ndown=ndown, ntop=ntop,
                coarse=coarse)
        else:
            vcycle(u, q1, r, u1, s, geometry=geometry, bc=pbc, lpldisc=lpl, nup=nup, ndown=ndown, ntop=ntop,
                   coarse=coarse)


def resnorm(r, b, u, s, tree=None, conn=None, geometry=None, bc=None, lpldisc=None, *,
            coarse=COARSE_GS) -> float:
    """resnorm! (src/multigrid.jl:391-417); with LaplacianDiscretization(D, 4)
    r receives the residual."""
    op = make_op(geometry, bc, lpldisc, coarse, any_order=True)
    out = C.c_double()
    _check(lib().srmg_resnorm(u.ctx.h, r.id, b.id, u.id, float(s), C.byref(op), C.byref(out)))
    return out.value


GUESS_SMOOTH = 4  # level-1 sweeps after a learned first iterate (default)


def _opts(ndown, ntop, nup, omega, tol, max_cycles, use_graph, inner_cycles=1, guess=0, guess_smooth=GUESS_SMOOTH,
          first_fmg=False, inner_first=0):
    o = _SolveOpts()
    o.ndown, o.ntop, o.nup = int(ndown), int(ntop), int(nup)
    o.omega, o.tol, o.max_cycles, o.use_graph = float(omega), float(tol), int(max_cycles), int(use_graph)
    o.inner_cycles = int(inner_cycles)
    o.guess = int(guess)
    o.guess_smooth = int(guess_smooth)
    o.first_fmg = int(bool(first_fmg))
    o.inner_first = int(inner_first)
    return o


def solve(u, b, r, u1, s, geometry, bc, lpldisc, *, tol=1e-8, max_cycles=50, ndown=2, ntop=4,
          nup=2, omega=1.0, coarse=COARSE_EXACT, use_graph=True, guess=GUESS_NONE, guess_smooth=GUESS_SMOOTH,
          first_fmg=False):
    """Pre-fill ghosts, then V-cycles until resnorm(u) <= tol * resnorm(0)
    (first_fmg: the first cycle is fmg!). Returns (cycles, relative residual
    history). LaplacianDiscretization(D, 4) needs coarse=COARSE_GS (the
    reference's level-1 sweeps)."""
    op = make_op(geometry, bc, lpldisc, coarse, any_order=True)
    o = _opts(ndown, ntop, nup, omega, tol, max_cycles, use_graph, 1, guess, guess_smooth, first_fmg)
    hist = np.zeros(max_cycles + 1)
    n = C.c_int()
    _check(lib().srmg_solve(u.ctx.h, u.id, b.id, r.id, u1.id, float(s), C.byref(op), C.byref(o),
                            C.byref(n), _p(hist)))
    return n.value, hist[: n.value + 1].copy()


def solve_mixed(u, b, e, be, re, ue, s, geometry, bc, lpldisc, *, tol=1e-8, max_cycles=50,
                ndown=2, ntop=4, nup=2, omega=1.0, coarse=COARSE_EXACT, use_graph=True, inner_cycles=1,
                guess=GUESS_NONE, guess_smooth=GUESS_SMOOTH, first_fmg=False, inner_first=0):
    """FP64 defect correction around FP32 V-cycles (u, b FP64; e, be, re, ue FP32;
    first_fmg: the first correction runs fmg! on e before its V-cycles;
    inner_first: V-cycles of the first correction, 0 for inner_cycles).
    Returns (outer steps, relative residual history per outer step).
    LaplacianDiscretization(D, 4) needs coarse=COARSE_GS."""
    op = make_op(geometry, bc, lpldisc, coarse, any_order=True)
    o = _opts(ndown, ntop, nup, omega, tol, max_cycles, use_graph, inner_cycles, guess, guess_smooth, first_fmg,
              inner_first)
    hist = np.zeros(max_cycles + 1)
    n = C.c_int()
    _check(lib().srmg_solve_mixed(u.ctx.h, u.id, b.id, e.id, be.id, re.id, ue.id, float(s),
                                  C.byref(op), C.byref(o), C.byref(n), _p(hist)))
    return n.value, hist[: n.value + 1].copy()


# ---- forests: several independent problems per launch ----

def concat_fields(fields):
    """Host fields of several trees concatenated along the block axis (the
    host layout of a forest)."""
    return np.asfortranarray(np.concatenate([f.u if hasattr(f, "u") else f for f in fields], axis=-1))


def split_field(arr, nblocks):
    """Inverse of :func:`concat_fields`: per-tree views."""
    out, o = [], 0
    for n in nblocks:
        out.append(arr[..., o:o + n])
        o += n
    return out


def _ops(geometry, bc, discs, coarse):
    arr = (_Op * len(discs))()
    for i, d in enumerate(discs):
        arr[i] = make_op(geometry, bc, d, coarse)
    return arr


def vcycle_batched(u, b, r, u1, s, geometry, bc, discs, *, lmax=0, ndown=3, ntop=5, nup=3, omega=1.0,
                   coarse=COARSE_GS):
    """vcycle! on every tree of a forest: tree t with scale s[t] and discretization discs[t]."""
    sv = np.ascontiguousarray(np.asarray(s, np.float64))
    ops = _ops(geometry, bc, discs, coarse)
    _check(lib().srmg_vcycle_batched(u.ctx.h, u.id, b.id, r.id, u1.id, _p(sv), ops, int(lmax), int(ndown),
                                     int(ntop), int(nup), float(omega)))


def resnorm_batched(r, b, u, s, geometry, bc, discs, *, coarse=COARSE_GS):
    sv = np.ascontiguousarray(np.asarray(s, np.float64))
    ops = _ops(geometry, bc, discs, coarse)
    out = np.zeros(len(discs))
    _check(lib().srmg_resnorm_batched(u.ctx.h, r.id, b.id, u.id, _p(sv), ops, _p(out)))
    return out


def solve_batched(u, b, r, u1, s, geometry, bc, discs, *, tol=1e-8, max_cycles=50, ndown=2, ntop=4, nup=2,
                  omega=1.0, coarse=COARSE_EXACT, use_graph=True, guess=GUESS_NONE, guess_smooth=GUESS_SMOOTH,
          first_fmg=False):
    """Solve every tree of a forest to tol; returns (cycles, history[ntrees, cycles + 1])."""
    sv = np.ascontiguousarray(np.asarray(s, np.float64))
    ops = _ops(geometry, bc, discs, coarse)
    o = _opts(ndown, ntop, nup, omega, tol, max_cycles, use_graph, 1, guess, guess_smooth, first_fmg)
    hist = np.zeros((len(discs), max_cycles + 1))
    n = C.c_int()
    _check(lib().srmg_solve_batched(u.ctx.h, u.id, b.id, r.id, u1.id, _p(sv), ops, C.byref(o), C.byref(n),
                                    _p(hist)))
    return n.value, hist[:, : n.value + 1].copy()


def solve_mixed_batched(u, b, e, be, re, ue, s, geometry, bc, discs, *, tol=1e-8, max_cycles=50, ndown=2,
                        ntop=4, nup=2, omega=1.0, coarse=COARSE_EXACT, use_graph=True, inner_cycles=1,
                        guess=GUESS_NONE, guess_smooth=GUESS_SMOOTH, first_fmg=False, inner_first=0):
    """FP64 defect correction around FP32 V-cycles on every tree of a forest."""
    sv = np.ascontiguousarray(np.asarray(s, np.float64))
    ops = _ops(geometry, bc, discs, coarse)
    o = _opts(ndown, ntop, nup, omega, tol, max_cycles, use_graph, inner_cycles, guess, guess_smooth, first_fmg,
              inner_first)
    hist = np.zeros((len(discs), max_cycles + 1))
    n = C.c_int()
    _check(lib().srmg_solve_mixed_batched(u.ctx.h, u.id, b.id, e.id, be.id, re.id, ue.id, _p(sv), ops,
                                          C.byref(o), C.byref(n), _p(hist)))
    return n.value, hist[:, : n.value + 1].copy()


# ---- electric field ----

def electric_field(u: DeviceField, eabs: DeviceField, h: float, e0, bc, G: int = 2, with_faces: bool = True):
    """electric_field!(tree, e, eabs, u, h, t, eb) (src/multigrid.jl:421-451)
    for a constant external field e0: |E| into the device field eabs; returns
    E in the reference's VectorBlockField layout (or None)."""
    c = u.ctx
    e = np.ascontiguousarray(np.asarray(e0, np.float64))
    b = _bc_arr(bc)
    out = None
    if with_faces:
        S = c.M + 2 * G + 1
        out = np.zeros((S,) * c.D + (c.D, c.nblocks), order="F")
    _check(lib().srmg_electric_field(c.h, u.id, eabs.id, float(h), _p(e), _p(b),
                                     None if out is None else out.ctypes.data_as(C.c_void_p), int(G)))
    return out


# ---- learned first iterate ----

def _ptr(x) -> int:
    """Device address of a torch tensor or an integer address."""
    return int(x.data_ptr()) if hasattr(x, "data_ptr") else int(x)


def grid_solve(ctx: Context, D: int, n, s: float, geometry, bc, disc, f, u, ngrid: int, noalloc: bool = False):
    """Exact FP64 solve of L_1 u = f on ngrid stacked uniform level-1 grids
    (device buffers, x fastest). noalloc: srmg_grid_solve_noalloc, which
    allocates no device memory (its grids are limited by the library's fixed
    scratch array)."""
    op = make_op(geometry, bc, disc, COARSE_EXACT)
    nn = (C.c_int * 3)(*(list(n) + [1] * (3 - len(n))))
    fn = lib().srmg_grid_solve_noalloc if noalloc else lib().srmg_grid_solve
    _check(fn(ctx.h, int(D), nn, float(s), C.byref(op), C.c_void_p(_ptr(f)), C.c_void_p(_ptr(u)), int(ngrid)))


def set_guess(ctx: Context, net: Optional["UNet"]):
    _check(lib().srmg_set_guess(ctx.h, None if net is None else net.h))
    ctx._guess = net  # keep the network alive while attached


def guess_fallbacks(ctx: Context) -> int:
    n = C.c_int()
    _check(lib().srmg_guess_fallbacks(ctx.h, C.byref(n)))
    return n.value


class UNet:
    """The network of the learned first iterate (srmg_unet_* in the C ABI).

    Data arguments are device buffers (torch tensors or addresses) of FP32
    stacked grids, x fastest."""

    def __init__(self, D: int, n, levels: int = 4, widths=(8, 16, 32, 64), seed: int = 0):
        self.D, self.n, self.levels, self.widths = int(D), tuple(int(x) for x in n), int(levels), tuple(widths)
        h = C.c_void_p()
        nn = (C.c_int * 3)(*(list(self.n) + [1] * (3 - len(self.n))))
        ww = (C.c_int * len(self.widths))(*self.widths)
        _check(lib().srmg_unet_create(C.byref(h), self.D, nn, self.levels, ww, C.c_uint64(seed)))
        self.h = h

    @property
    def nparams(self) -> int:
        n = C.c_int64()
        _check(lib().srmg_unet_param_count(self.h, C.byref(n)))
        return n.value

    def get_params(self) -> np.ndarray:
        out = np.zeros(self.nparams, np.float32)
        _check(lib().srmg_unet_get_params(self.h, _p(out)))
        return out

    def set_params(self, p):
        p = np.ascontiguousarray(np.asarray(p, np.float32))
        if p.size != self.nparams:
            raise SrmgError(f"expected {self.nparams} parameters, got {p.size}")
        _check(lib().srmg_unet_set_params(self.h, _p(p)))

    @property
    def output_scale(self) -> float:
        c = C.c_double()
        _check(lib().srmg_unet_get_output_scale(self.h, C.byref(c)))
        return c.value

    @output_scale.setter
    def output_scale(self, c: float):
        _check(lib().srmg_unet_set_output_scale(self.h, float(c)))

    def set_problem(self, geometry, bc, disc):
        op = make_op(geometry, bc, disc, COARSE_EXACT)
        _check(lib().srmg_unet_set_problem(self.h, C.byref(op)))
        self.problem = (int(op.geometry), np.asarray(bc, np.int8).ravel().copy(), int(op.disc), float(op.k2))

    def forward(self, f, u, batch: int):
        _check(lib().srmg_unet_forward(self.h, C.c_void_p(_ptr(f)), C.c_void_p(_ptr(u)), int(batch)))

    def loss_grad(self, f, u, batch: int):
        loss = C.c_double()
        g = np.zeros(self.nparams, np.float32)
        _check(lib().srmg_unet_loss_grad(self.h, C.c_void_p(_ptr(f)), C.c_void_p(_ptr(u)), int(batch),
                                         C.byref(loss), _p(g)))
        return loss.value, g

    def train_step(self, f, u, batch: int, lr: float, want_loss: bool = False):
        loss = C.c_double()
        _check(lib().srmg_unet_train_step(self.h, C.c_void_p(_ptr(f)), C.c_void_p(_ptr(u)), int(batch),
                                          C.c_float(lr), C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def synchronize(self):
        _check(lib().srmg_unet_synchronize(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().srmg_unet_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
