"""ctypes binding of the CPU oracle (liboracle.so built from srmg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, never by the library itself. Every function
mirrors a reference routine named in srmg_oracle.c and works in place on host
fields in the reference's layout (``ScalarBlockField.u``, Fortran order).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

GEOM = {"cartesian": 0, "cyl1": 1, "cyl2": 2}


def build(force: bool = False) -> str:
    path = os.path.join(_HERE, "liboracle.so")
    if force or not os.path.exists(path):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    return path


def lib():
    global _LIB
    if _LIB is None:
        _LIB = C.CDLL(build())
        P = C.c_void_p
        i64, i32, dbl = C.c_int64, C.c_int, C.c_double
        _LIB.ora_tree_create.restype = P
        _LIB.ora_tree_create.argtypes = [i32, i32, i32, i32, i64] + [P] * 10
        _LIB.ora_tree_destroy.argtypes = [P]
        for name, args in {
            "ora_fill_ghost_copy": [P, P, i32],
            "ora_fill_ghost_bnd": [P, P, i32, P],
            "ora_fill_ghost_interp": [P, P, i32],
            "ora_fill_ghost": [P, P, i32, P],
            "ora_fill_ghost_all": [P, P, P],
            "ora_gs_level": [P, P, P, dbl, dbl, i32, P, i32],
            "ora_gs_iter": [P, P, P, dbl, dbl, i32, i32, P],
            "ora_residual_level": [P, P, P, P, dbl, i32, P],
            "ora_restrict_level": [P, P, i32],
            "ora_postrestrict_level": [P, P, P, P, dbl, i32, P],
            "ora_copy_level": [P, P, P, i32],
            "ora_diff_level": [P, P, P, P, i32],
            "ora_interp_add_level": [P, P, P, i32],
            "ora_vcycle": [P, P, P, P, P, dbl, P, i32, i32, i32, i32, dbl, i32],
            "ora_fmg": [P, P, P, P, P, dbl, P, i32, i32, i32, dbl, i32],
        }.items():
            getattr(_LIB, name).argtypes = args
            getattr(_LIB, name).restype = None
        _LIB.ora_resnorm.argtypes = [P, P, P, P, dbl, P]
        _LIB.ora_resnorm.restype = dbl
        _LIB.ora_coarse_solve.argtypes = [P, P, P, dbl, P]
        _LIB.ora_coarse_solve.restype = i32
        _LIB.ora_set_threads.argtypes = [i32]
        _LIB.ora_set_threads.restype = None
        _LIB.ora_max_threads.argtypes = []
        _LIB.ora_max_threads.restype = i32
        _LIB.ora_axis_decomp.argtypes = [i32, i32, dbl, dbl, P, P, P]
        _LIB.ora_axis_decomp.restype = i32
        _LIB.ora_axis_decomp_jacobi.argtypes = [i32, i32, dbl, dbl, P, P, P]
        _LIB.ora_axis_decomp_jacobi.restype = None
        _LIB.ora_electric_field.argtypes = [P, P, P, P, dbl, P]
        _LIB.ora_electric_field.restype = None
        for name, args in {
            "ora_freebc_getouter": [P, P, P, i32, i32, dbl, P],
            "ora_dst_matrix": [P, i64, i32],
            "ora_freebc_setouter": [P, P, P, i32, i32, dbl, P],
            "ora_freebc": [P, P, P, dbl, dbl, i32, i32, P, P, i64, P],
            "ora_fill_ghost_free": [P, P, P, i32, i32, i32, P],
            "ora_rhs_level": [P, P, P, i32, P, i32],
            "ora_gs4_level": [P, P, P, C.c_double, C.c_double, i32, P, i32, i32],
        }.items():
            getattr(_LIB, name).argtypes = args
            getattr(_LIB, name).restype = None
    return _LIB


def set_threads(n: int) -> None:
    """OpenMP threads of the oracle's level loops."""
    lib().ora_set_threads(int(n))


def max_threads() -> int:
    return int(lib().ora_max_threads())


def axis_decomp(n: int, radial: bool, alow: float, ahigh: float, method: str = "ql"):
    """The level-1 1D operator of one axis as T = P diag(lam) Q: the implicit
    QL eigensolver of the exact coarse solve (method "ql") or the cyclic
    Jacobi cross-check ("jacobi")."""
    lam = np.zeros(n)
    P = np.zeros((n, n))
    Q = np.zeros((n, n))
    if method == "ql":
        if lib().ora_axis_decomp(int(n), int(bool(radial)), float(alow), float(ahigh), _ptr(lam), _ptr(P), _ptr(Q)):
            raise RuntimeError("QL iteration did not converge")
    else:
        lib().ora_axis_decomp_jacobi(int(n), int(bool(radial)), float(alow), float(ahigh), _ptr(lam), _ptr(P),
                                     _ptr(Q))
    return lam, P, Q


def dst_matrix(n: int, kind: int) -> np.ndarray:
    """The n x n matrix of FFTW's RODFT10 (kind 2) or RODFT01 (kind 3)."""
    S = np.zeros((n, n))
    lib().ora_dst_matrix(_ptr(S), int(n), int(kind))
    return S


# ---- applydelta!'s field part (numpy restatement) ----------------------------

def interp_block(dest: np.ndarray, src: np.ndarray, sub, M: int, G: int):
    """interp!(dest, validindices, src, subblockindices(sub)) with InterpLinear
    and func = _first (src/interp.jl:7-11, 23-60, 88-91): the valid cells of
    block array ``dest`` ((M+2G)^D) from the subblock ``sub`` (1-based) of
    block array ``src``, whose depth-1 ghosts must be filled. Each fine cell is
    w0 src[Js] + sum over dimensions j (in order) of w1 src[Js +- e_j], the
    sign being that of the fine cell's offset."""
    D = src.ndim
    h = M // 2
    w0, w1 = (0.5, 0.25) if D == 2 else (0.25, 0.25)
    for par in np.ndindex(*(2,) * D):
        js = tuple(slice(G + (sub[d] - 1) * h, G + (sub[d] - 1) * h + h) for d in range(D))
        f = w0 * src[js]
        for j in range(D):
            sh = 2 * par[j] - 1
            jj = tuple(slice(js[d].start + (sh if d == j else 0), js[d].stop + (sh if d == j else 0))
                       for d in range(D))
            f = f + w1 * src[jj]
        dest[tuple(slice(G + par[d], G + M, 2) for d in range(D))] = f


def restrict_block(dest: np.ndarray, sub, src: np.ndarray, M: int, G: int):
    """restrict!(dest, subblockindices(sub), src, validindices) (src/restrict.jl:5-26):
    each cell of the subblock ``sub`` of ``dest`` = the sum of its 2^D children
    of ``src`` (first index fastest, from zero) divided by 2^D."""
    D = src.ndim
    h = M // 2
    f = np.zeros((h,) * D, dtype=src.dtype)
    for par in [tuple(reversed(t)) for t in np.ndindex(*(2,) * D)]:
        f = f + src[tuple(slice(G + par[d], G + M, 2) for d in range(D))]
    dest[tuple(slice(G + (sub[d] - 1) * h, G + (sub[d] - 1) * h + h) for d in range(D))] = f / src.dtype.type(1 << D)


def applydelta_fields(u: np.ndarray, actions, nblocks: int, M: int, G: int, interp: bool = True,
                      restrict: bool = True) -> np.ndarray:
    """The field part of applydelta! (src/refinement.jl:339-370; interp! and
    restrict! of src/streamer.jl:191-221) on a host field ``u`` whose ghosts
    were filled on the old tree: a new array of ``nblocks`` blocks (ids are
    stable; new ids start at zero), then the actions of
    ``srmg.tree.applydelta`` in its order."""
    out = np.zeros(u.shape[:-1] + (nblocks,), dtype=u.dtype, order="F")
    out[..., :u.shape[-1]] = u
    for kind, a, b, sub in actions:
        if kind == "restrict" and restrict:
            restrict_block(out[..., a - 1], sub, out[..., b - 1], M, G)
        elif kind == "interp" and interp:
            interp_block(out[..., a - 1], out[..., b - 1], sub, M, G)
    return out


class _Op(C.Structure):
    _fields_ = [("geom", C.c_int), ("helm", C.c_int), ("k2", C.c_double),
                ("bc", C.c_int8 * 27), ("order", C.c_int)]


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """Oracle bound to one tree. ``tree``/``conn`` are the host structures of
    ``srmg.tree``; fields are ``ScalarBlockField`` objects (or their ``.u``)."""

    def __init__(self, tree, conn, M: int, G: int, geom: int, bc, helm: bool = False,
                 k2: float = 0.0, order: int = 2):
        from_tree = _flatten(tree, conn)
        self.D, self.M, self.G = tree.D, M, G
        self.nlevels = len(tree)
        self.nblocks = tree.nblocks()
        self._keep = from_tree
        L = lib()
        self.h = L.ora_tree_create(self.D, M, G, self.nlevels, self.nblocks,
                                   *[_ptr(a) for a in from_tree])
        self.op = _Op()
        self.op.geom = int(geom)
        self.op.helm = int(bool(helm))
        self.op.k2 = float(k2)
        self.op.order = int(order)  # 4: LaplacianDiscretization{D,4}, Cartesian (box-stencil records)
        bc = np.asarray(bc, np.int8).ravel()
        for i in range(len(bc)):
            self.op.bc[i] = int(bc[i])
        self._bc = np.ascontiguousarray(bc, dtype=np.int8)

    @classmethod
    def for_problem(cls, p, G=None):
        return cls(p.tree, p.conn, p.M, p.b.G if G is None else G, p.geom, p.bc, p.helm, p.k2)

    def __del__(self):
        try:
            lib().ora_tree_destroy(self.h)
        except Exception:
            pass

    @staticmethod
    def _a(f):
        a = f.u if hasattr(f, "u") else f
        assert a.dtype == np.float64 and a.flags.f_contiguous
        return a

    def _opp(self):
        return C.byref(self.op)

    # ghost fills -----------------------------------------------------------
    def fill_ghost(self, u, level):
        lib().ora_fill_ghost(self.h, _ptr(self._a(u)), level, _ptr(self._bc))

    def fill_ghost_all(self, u):
        lib().ora_fill_ghost_all(self.h, _ptr(self._a(u)), _ptr(self._bc))

    def fill_ghost_copy(self, u, level):
        lib().ora_fill_ghost_copy(self.h, _ptr(self._a(u)), level)

    def fill_ghost_bnd(self, u, level):
        lib().ora_fill_ghost_bnd(self.h, _ptr(self._a(u)), level, _ptr(self._bc))

    def fill_ghost_interp(self, u, level):
        lib().ora_fill_ghost_interp(self.h, _ptr(self._a(u)), level)

    # level operations ------------------------------------------------------
    def gs_iter(self, u, b, omega, s, n, level):
        lib().ora_gs_iter(self.h, _ptr(self._a(u)), _ptr(self._a(b)), omega, s, n, level, self._opp())

    def gs_level(self, u, b, omega, s, level, P):
        lib().ora_gs_level(self.h, _ptr(self._a(u)), _ptr(self._a(b)), omega, s, level, self._opp(), P)

    def residual_level(self, r, u, b, s, level):
        lib().ora_residual_level(self.h, _ptr(self._a(r)), _ptr(self._a(u)), _ptr(self._a(b)),
                                 s, level, self._opp())

    def gs4_level(self, u, b, omega, s, level, P1, P2):
        """one FourColors{P1,P2} pass of gauss_seidel! over a level (order 4)."""
        lib().ora_gs4_level(self.h, _ptr(self._a(u)), _ptr(self._a(b)), float(omega), float(s), level,
                            self._opp(), int(P1), int(P2))

    def rhs_level(self, b1, b, level, order=2):
        """rhs_level!(b1, b, tree[level], lpl, geom) for a discretization of
        order 2 or 4 (b's ghosts of this level filled)."""
        lib().ora_rhs_level(self.h, _ptr(self._a(b1)), _ptr(self._a(b)), level, self._opp(), int(order))

    def rhs(self, b1, b, order=2):
        """derivs!'s loop (src/streamer.jl:314-318): fill_ghost!(b, l), then
        rhs_level!(b1, b, l), level by level."""
        for l in range(1, self.nlevels + 1):
            self.fill_ghost(b, l)
            self.rhs_level(b1, b, l, order)

    def restrict_full(self, u):
        """restrict_full!(u, conn) (src/restrict.jl:43-47): finest level first."""
        for l in range(self.nlevels, 1, -1):
            self.restrict_level(u, l)

    def restrict_level(self, u, level):
        lib().ora_restrict_level(self.h, _ptr(self._a(u)), level)

    def postrestrict_level(self, b, u, r, s, level):
        lib().ora_postrestrict_level(self.h, _ptr(self._a(b)), _ptr(self._a(u)), _ptr(self._a(r)),
                                     s, level, self._opp())

    def copy_level(self, dst, src, level):
        lib().ora_copy_level(self.h, _ptr(self._a(dst)), _ptr(self._a(src)), level)

    def diff_level(self, dst, a1, a2, level):
        lib().ora_diff_level(self.h, _ptr(self._a(dst)), _ptr(self._a(a1)), _ptr(self._a(a2)), level)

    def interp_add_level(self, u, u1, level):
        lib().ora_interp_add_level(self.h, _ptr(self._a(u)), _ptr(self._a(u1)), level)

    # solvers ---------------------------------------------------------------
    def vcycle(self, u, b, r, u1, s, lmax=0, ndown=3, ntop=5, nup=3, omega=1.0, coarse_exact=False):
        lib().ora_vcycle(self.h, _ptr(self._a(u)), _ptr(self._a(b)), _ptr(self._a(r)),
                         _ptr(self._a(u1)), s, self._opp(), lmax, ndown, ntop, nup, omega,
                         int(bool(coarse_exact)))

    def fmg(self, u, b, r, u1, s, ndown=3, ntop=5, nup=3, omega=1.25, coarse_exact=False):
        lib().ora_fmg(self.h, _ptr(self._a(u)), _ptr(self._a(b)), _ptr(self._a(r)),
                      _ptr(self._a(u1)), s, self._opp(), ndown, ntop, nup, omega,
                      int(bool(coarse_exact)))

    def resnorm(self, r, b, u, s) -> float:
        return lib().ora_resnorm(self.h, _ptr(self._a(r)), _ptr(self._a(b)), _ptr(self._a(u)),
                                 s, self._opp())

    def electric_field(self, u, h, e0):
        """electric_field! for a constant external field: (E in the
        VectorBlockField layout, |E| as a host field array). u's ghosts must be
        filled (fill_ghost_all)."""
        a = self._a(u)
        S = self.M + 2 * self.G + 1
        f = np.zeros((S,) * self.D + (self.D, self.nblocks), order="F")
        fabs = np.zeros_like(a)
        e = np.ascontiguousarray(np.asarray(e0, np.float64))
        lib().ora_electric_field(self.h, _ptr(f), _ptr(fabs), _ptr(a), float(h), _ptr(e))
        return f, fabs

    # free boundary conditions (src/freebc.jl) -----------------------------
    def freebc_getouter(self, u, dim, level, h, rootsize, n):
        u1 = np.zeros(n)
        root = np.ascontiguousarray(np.asarray(rootsize, np.int64))
        lib().ora_freebc_getouter(self.h, _ptr(u1), _ptr(self._a(u)), int(dim), int(level), float(h), _ptr(root))
        return u1

    def freebc_setouter(self, q, b, dim, level, s, rootsize):
        root = np.ascontiguousarray(np.asarray(rootsize, np.int64))
        bb = np.ascontiguousarray(np.asarray(b, np.float64))
        lib().ora_freebc_setouter(self.h, _ptr(self._a(q)), _ptr(bb), int(dim), int(level), float(s), _ptr(root))

    def freebc(self, q, u, h, s, dim, level, rootsize, f):
        """freebc!(q, u, h, s, tree, conn, conf); returns the boundary values b."""
        root = np.ascontiguousarray(np.asarray(rootsize, np.int64))
        ff = np.ascontiguousarray(np.asarray(f, np.float64))
        b = np.zeros(len(ff))
        lib().ora_freebc(self.h, _ptr(self._a(q)), _ptr(self._a(u)), float(h), float(s), int(dim), int(level),
                         _ptr(root), _ptr(ff), len(ff), _ptr(b))
        return b

    def fill_ghost_free(self, u, b, dim, level, l, rootsize):
        root = np.ascontiguousarray(np.asarray(rootsize, np.int64))
        bb = np.ascontiguousarray(np.asarray(b, np.float64))
        lib().ora_fill_ghost_free(self.h, _ptr(self._a(u)), _ptr(bb), int(dim), int(level), int(l), _ptr(root))

    def coarse_solve(self, u, b, s) -> int:
        return lib().ora_coarse_solve(self.h, _ptr(self._a(u)), _ptr(self._a(b)), s, self._opp())


def _flatten(tree, conn):
    """pairs and the four record lists as flat int64 arrays + level offsets."""
    from_pairs = []
    off = [0]
    for layer in tree:
        for c, blk in layer.pairs:
            from_pairs.append((*c, blk))
        off.append(len(from_pairs))
    D = tree.D
    pairs = np.ascontiguousarray(np.array(from_pairs, np.int64).reshape(-1, D + 1))
    out = [pairs, np.array(off, np.int64)]
    for lists, w in ((conn.neighbor, 2 + D), (conn.boundary, 2 + 2 * D),
                     (conn.refboundary, 2 + 2 * D), (conn.child, 2 + D)):
        o = [0]
        for a in lists:
            o.append(o[-1] + len(a))
        arr = (np.concatenate([a.reshape(-1, w) for a in lists]) if o[-1]
               else np.zeros((1, w), np.int64))
        out += [np.ascontiguousarray(arr.astype(np.int64)), np.array(o, np.int64)]
    return out
