"""Brute-force pin of the oracle: a pure-Python transliteration of the
reference's loops, run on tiny trees and compared bit for bit with the oracle.

Every routine below follows the reference's Julia code structure: 1-based
ranges, the stencil literal iterated column-major with zeros skipped
(src/laplacian.jl:25-49), the ranges of ghostindices / mirrorghost /
subblockbnd (src/blockfield.jl:407-541), the loop nests of restrict!/interp!
(src/restrict.jl:5-25, src/interp.jl:102-141) and the call sequence of
vcycle!/fmg! (src/multigrid.jl:17-106). `muladd` is evaluated exactly and
rounded once (src/multigrid.jl:176-177), via rational arithmetic. It shares no
code with the oracle, so a mistranslated range, sign, colour or call order in
either shows up as a mismatch.
"""

import itertools
from fractions import Fraction

import numpy as np
import pytest

from oracle import Oracle
from srmg import configs


def muladd(a, b, c):
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def cartesian(ranges):
    """CartesianIndices over 1-based inclusive (lo, hi) ranges, first index fastest."""
    for t in itertools.product(*[range(lo, hi + 1) for lo, hi in reversed(ranges)]):
        yield tuple(reversed(t))


class Ref:
    denom = 1  # the @stencil denominator of the order-2 operators

    def __init__(self, p, conn=None):
        self.p = p
        self.D, self.M, self.G = p.D, p.M, p.b.G
        self.tree, self.conn = p.tree, (p.conn if conn is None else conn)
        self.bc = p.bc.reshape((3,) * self.D, order="F")

    # -- fields: Julia-layout arrays indexed with 1-based tuples ------------
    def get(self, f, I, blk):
        return f.u[tuple(i - 1 for i in I) + (blk - 1,)]

    def set(self, f, I, blk, v):
        f.u[tuple(i - 1 for i in I) + (blk - 1,)] = v

    # -- src/laplacian.jl, src/helmholtz.jl ---------------------------------
    def literal(self, J, s):
        D, p = self.D, self.p
        cc = (-4 if D == 2 else -6)
        if p.helm:
            cc = cc - s * p.k2
        if D == 2:
            if p.geom == 0:
                return [[0, 1, 0], [1, cc, 1], [0, 1, 0]]
            i = J[0] if p.geom == 1 else J[1]
            wm, wp = (2 * i - 2) / (2 * i - 1), 2 * i / (2 * i - 1)
            if p.geom == 1:
                return [[0, wm, 0], [1, cc, 1], [0, wp, 0]]
            return [[0, 1, 0], [wm, cc, wp], [0, 1, 0]]
        a1 = [[0, 0, 0], [0, 1, 0], [0, 0, 0]]
        a2 = [[0, 1, 0], [1, cc, 1], [0, 1, 0]]
        # arr[i1][i2][i3]: the third index selects a1 / a2 / a1
        return [[[(a1, a2, a1)[k][i][j] for k in range(3)] for j in range(3)] for i in range(3)]

    def applystencil(self, u, blk, I, J, s):
        arr = self.literal(J, s)
        v = 0.0
        for S1 in cartesian([(1, 3)] * self.D):
            c = arr
            for k in S1:
                c = c[k - 1]
            if c == 0:
                continue
            I1 = tuple(I[d] + S1[d] - 2 for d in range(self.D))
            v += c * self.get(u, I1, blk)
        return v / self.denom

    def diagelm(self, s):
        c = -4 if self.D == 2 else -6
        return c - s * self.p.k2 if self.p.helm else c

    def J(self, I, blkpos):
        gbl0 = tuple(self.M * (b - 1) + 1 for b in blkpos)
        return tuple(I[d] - self.G + gbl0[d] - 1 for d in range(self.D))

    # -- src/multigrid.jl ----------------------------------------------------
    def gauss_seidel(self, u, b, omega, s, blkpos, blk, P):
        G, M, D = self.G, self.M, self.D
        ranges = [list(range(G + 1, G + M + 1, 2))] + [list(range(G + 1, G + M + 1))] * (D - 1)
        for t in itertools.product(*reversed(ranges)):
            i = tuple(reversed(t))
            p = P
            for d in range(1, D):
                p ^= i[d] & 1
            I = (i[0] + p,) + i[1:]
            J = self.J(I, blkpos)
            c = self.diagelm(s)
            Lu = self.applystencil(u, blk, I, J, s)
            v = muladd(s, self.get(b, I, blk), Lu)
            self.set(u, I, blk, muladd(-(omega / c), v, self.get(u, I, blk)))

    def gauss_seidel_iter(self, u, b, omega, s, n, l):
        for _ in range(n):
            for P in (0, 1):
                for c, blk in self.tree[l - 1].pairs:
                    self.gauss_seidel(u, b, omega, s, c, blk, P)
                self.fill_ghost_copy(u, l)
                self.fill_ghost_bnd(u, l)

    def residual(self, r, u, b, s, blkpos, blk, inds):
        for I in inds:
            J = self.J(I, blkpos)
            Lu = self.applystencil(u, blk, I, J, s)
            self.set(r, I, blk, -(self.get(b, I, blk) + Lu / s))

    def residual_level(self, r, u, b, s, l):
        G, M = self.G, self.M
        for c, blk in self.tree[l - 1].pairs:
            self.residual(r, u, b, s, c, blk, list(cartesian([(G + 1, G + M)] * self.D)))

    def residual_postrestrict_level(self, bout, u, r, s, l):
        for rec in self.conn.child[l - 1]:
            coarse, sb = int(rec[1]), rec[2:]
            blkpos = self.tree[l - 2].coord[coarse]
            self.residual(bout, u, r, s, blkpos, coarse, list(cartesian(self.subblockindices(sb))))

    # -- src/blockfield.jl ---------------------------------------------------
    def subblockindices(self, sb):
        G, mh = self.G, self.M // 2
        return [(G + 1 + (int(sb[d]) - 1) * mh, G + int(sb[d]) * mh) for d in range(self.D)]

    def subblockbnd(self, sb, face):
        ghalf = self.G // 2
        out = []
        for d, (lo, hi) in enumerate(self.subblockindices(sb)):
            if face[d] == -1:
                out.append((lo, lo + ghalf - 1))
            elif face[d] == 1:
                out.append((hi - ghalf + 1, hi))
            else:
                out.append((lo, hi))
        return out

    def ghostindices(self, face):
        G, M = self.G, self.M
        out = []
        for f in face:
            if f == 0:
                out.append((G + 1, G + M))
            elif f == -1:
                out.append((1, G))
            else:
                out.append((M + G + 1, M + 2 * G))
        return out

    def mirror(self, i, d):
        G, M = self.G, self.M
        if d == -1:
            return 2 * G - i + 1
        if d == 1:
            return 2 * (G + M) - i + 1
        return i

    # -- src/connectivity.jl -------------------------------------------------
    def fill_ghost_copy(self, u, l):
        G, M, D = self.G, self.M, self.D
        for rec in self.conn.neighbor[l - 1]:
            frm, to, f = int(rec[0]), int(rec[1]), rec[2:]
            for i in cartesian([(1, M) if f[d] == 0 else (1, G) for d in range(D)]):
                ifrom = tuple(G + i[d] if f[d] == 0 else (i[d] + G if f[d] == -1 else i[d] + M)
                              for d in range(D))
                ito = tuple(G + i[d] if f[d] == 0 else (i[d] if f[d] == 1 else i[d] + M + G)
                            for d in range(D))
                self.set(u, ito, to, self.get(u, ifrom, frm))

    def fill_ghost_bnd(self, u, l):
        D = self.D
        for rec in self.conn.boundary[l - 1]:
            blk, face, bnd = int(rec[0]), rec[1:1 + D], rec[1 + D:1 + 2 * D]
            a = int(self.bc[tuple(2 + int(x) - 1 for x in bnd)])
            for g in cartesian(self.ghostindices(face)):
                v = tuple(self.mirror(g[d], bnd[d]) for d in range(D))
                self.set(u, g, blk, a * self.get(u, v, blk))

    def interp(self, dst, dblk, idest, src, sblk, isrc, add):
        D = self.D
        w0, w1 = (1 / 2, 1 / 4) if D == 2 else (1 / 4, 1 / 4)
        for Is in cartesian([(1, hi - lo + 1) for lo, hi in isrc]):
            Id0 = tuple(2 * x - 1 for x in Is)
            Jd0 = tuple(idest[d][0] + Id0[d] - 1 for d in range(D))
            Js = tuple(isrc[d][0] + Is[d] - 1 for d in range(D))
            for i in cartesian([(0, 1)] * D):
                Jd = tuple(Jd0[d] + i[d] for d in range(D))
                f = w0 * self.get(src, Js, sblk)
                for j in range(D):
                    K = tuple(Js[d] + 2 * i[d] - 1 if d == j else Js[d] for d in range(D))
                    f += w1 * self.get(src, K, sblk)
                self.set(dst, Jd, dblk, f + self.get(dst, Jd, dblk) if add else f)

    def fill_ghost_interp(self, u, l):
        D = self.D
        for rec in self.conn.refboundary[l - 1]:
            fine, coarse, face, sb = int(rec[0]), int(rec[1]), rec[2:2 + D], rec[2 + D:]
            self.interp(u, fine, self.ghostindices(face), u, coarse,
                        self.subblockbnd(sb, [-int(x) for x in face]), False)

    def fill_ghost(self, u, l):
        self.fill_ghost_copy(u, l)
        self.fill_ghost_bnd(u, l)
        self.fill_ghost_interp(u, l)

    # -- src/restrict.jl, src/interp.jl, src/blockfield.jl:557-576 ----------
    def restrict_level(self, u, l):
        D, G, M = self.D, self.G, self.M
        k = 1 << D
        for rec in self.conn.child[l - 1]:
            fine, coarse, sb = int(rec[0]), int(rec[1]), rec[2:]
            idest = self.subblockindices(sb)
            for Id in cartesian([(1, hi - lo + 1) for lo, hi in idest]):
                Js0 = tuple(G + 1 + 2 * Id[d] - 2 for d in range(D))
                Jd = tuple(idest[d][0] + Id[d] - 1 for d in range(D))
                f = 0.0
                for i in cartesian([(0, 1)] * D):
                    f += self.get(u, tuple(Js0[d] + i[d] for d in range(D)), fine)
                self.set(u, Jd, coarse, f / k)

    def interp_add_level(self, u, u1, l):
        G, M = self.G, self.M
        for rec in self.conn.child[l - 1]:
            fine, coarse, sb = int(rec[0]), int(rec[1]), rec[2:]
            self.interp(u, fine, [(G + 1, G + M)] * self.D, u1, coarse, self.subblockindices(sb), True)

    def copyto(self, dst, src, l):
        for _, blk in self.tree[l - 1].pairs:
            dst.u[..., blk - 1] = src.u[..., blk - 1]

    def diffto(self, dst, a1, a2, l):
        for _, blk in self.tree[l - 1].pairs:
            dst.u[..., blk - 1] = a1.u[..., blk - 1] - a2.u[..., blk - 1]

    def vcycle(self, u, b, r, u1, s, lmax=0, ndown=3, ntop=5, nup=3, omega=1.0):
        if lmax == 0:
            lmax = len(self.tree)
        for l in range(lmax, 1, -1):
            sl = s / (1 << (l - 1)) ** 2
            self.gauss_seidel_iter(u, b, omega, sl, ndown, l)
            self.restrict_level(u, l)
            self.fill_ghost(u, l - 1)
            self.copyto(u1, u, l - 1)
            self.residual_level(r, u, b, sl, l)
            self.fill_ghost(r, l)
            self.restrict_level(r, l)
            self.fill_ghost(r, l - 1)
            self.residual_postrestrict_level(b, u, r, 4 * sl, l)
        self.gauss_seidel_iter(u, b, omega, s, ntop, 1)
        for l in range(2, lmax + 1):
            sl = s / (1 << (l - 1)) ** 2
            self.diffto(u1, u, u1, l - 1)
            self.fill_ghost(u1, l - 1)
            self.interp_add_level(u, u1, l)
            self.fill_ghost(u, l)
            self.gauss_seidel_iter(u, b, omega, sl, nup, l)
            self.fill_ghost(u, l)

    def fmg(self, u, b, r, u1, s, ndown=3, ntop=5, nup=3, omega=1.25):
        lmax = self.tree.finestlevel()
        for l in range(lmax, 1, -1):
            sl = s / (1 << (l - 1)) ** 2
            self.restrict_level(u, l)
            self.fill_ghost(u, l - 1)
            self.residual_level(r, u, b, sl, l)
            self.fill_ghost(r, l)
            self.restrict_level(r, l)
            self.fill_ghost(r, l - 1)
            self.residual_postrestrict_level(b, u, r, 4 * sl, l)
        for l in range(1, lmax + 1):
            self.copyto(u1, u, l)
            if l > 1:
                self.diffto(u1, u, u1, l - 1)
                self.fill_ghost(u1, l - 1)
                self.interp_add_level(u, u1, l)
            self.fill_ghost(u, l)
            self.vcycle(u, b, r, u1, s, lmax=l, ndown=ndown, ntop=ntop, nup=nup, omega=omega)


class Ref4(Ref):
    """LaplacianDiscretization{D,4}: the compact stencil literals of
    src/laplacian.jl:133-201 (2D Cartesian and CylindricalGeometry{1,2}) and
    305-340 (3D Cartesian) with their denominator 6, diagelm, the FourColors
    gauss_seidel! of src/multigrid.jl:206-232 (u -= ω (Lu + s b) / c, no
    muladd) and the four-pass gauss_seidel_iter! of src/multigrid.jl:268-295.
    Ghost fills run on the BoxStencil records."""
    denom = 6

    def literal(self, J, s):
        D, p = self.D, self.p
        if D == 3:
            a1 = [[0, 1, 0], [1, 2, 1], [0, 1, 0]]
            a2 = [[1, 2, 1], [2, -24, 2], [1, 2, 1]]
            return [[[(a1, a2, a1)[k][i][j] for k in range(3)] for j in range(3)] for i in range(3)]
        if p.geom == 0:
            return [[1, 4, 1], [4, -20, 4], [1, 4, 1]]
        i = J[0] if p.geom == 1 else J[1]
        # Int / Int in Julia: the exact ratio rounded once (Python's true division)
        lo = 2 * (i - 1) / (2 * i - 1)
        hi = 2 * i / (2 * i - 1)
        pm = 4 * (-2 + 9 * i - 16 * i ** 2 + 8 * i ** 3) / (2 * i - 1) ** 3
        cc = -16 * (1 - 5 * i + 5 * i ** 2) / (1 - 2 * i) ** 2
        pp = 4 * (1 + i - 8 * i ** 2 + 8 * i ** 3) / (2 * i - 1) ** 3
        if p.geom == 1:
            return [[lo, pm, lo], [4, cc, 4], [hi, pp, hi]]
        return [[lo, 4, hi], [pm, cc, pp], [lo, 4, hi]]

    def diagelm4(self, J):
        if self.D == 3:
            return -4
        if self.p.geom == 0:
            return -20 / 6
        i = J[0] if self.p.geom == 1 else J[1]
        return -16 * (1 - 5 * i + 5 * i ** 2) / (1 - 2 * i) ** 2 / 6

    def gauss_seidel4(self, u, b, omega, s, blkpos, blk, P1, P2):
        G, M, D = self.G, self.M, self.D
        ranges = [list(range(G + 1, G + M + 1, 2))] * 2 + [list(range(G + 1, G + M + 1))] * (D - 2)
        for t in itertools.product(*reversed(ranges)):
            I1 = tuple(reversed(t))
            p1, p2 = P1, P2
            if D == 3:
                p1 = (p1 + I1[2]) % 2
                p2 = (p2 + I1[2]) % 2
            I = (I1[0] + p1, I1[1] + p2) + I1[2:]
            J = self.J(I, blkpos)
            c = self.diagelm4(J)
            Lu = self.applystencil(u, blk, I, J, s)
            self.set(u, I, blk, self.get(u, I, blk) - omega * (Lu + s * self.get(b, I, blk)) / c)

    def gauss_seidel_iter(self, u, b, omega, s, n, l):
        for _ in range(n):
            for P1, P2 in ((0, 0), (0, 1), (1, 0), (1, 1)):
                for c, blk in self.tree[l - 1].pairs:
                    self.gauss_seidel4(u, b, omega, s, c, blk, P1, P2)
                self.fill_ghost_copy(u, l)
                self.fill_ghost_bnd(u, l)


def _tiny(D, geom, bcd, helm=False, k2=0.0):
    if D == 2:
        rootsize, M = (2, 3), 4

        def crit(l, c):
            return (c[:, 0] == 1) & (c[:, 1] == 2) if l == 1 else (c[:, 0] == 2) & (c[:, 1] == 3)
        levels = 3
    else:
        rootsize, M = (2, 2, 2), 4

        def crit(l, c):
            return np.all(c == np.array([1, 2, 1]), axis=1)
        levels = 2
    return configs.uniform_problem(D, rootsize, M, levels, geom, bcd, helm=helm, k2=k2, seed=11,
                                   refine=crit)


CASES = [
    (2, 1, ((1, -1), (-1, -1)), False, 0.0),
    (2, 2, ((-1, 1), (1, -1)), False, 0.0),
    (2, 0, ((1, 1), (-1, -1)), True, 3e3),
    (3, 0, ((1, 1), (1, -1), (-1, -1)), False, 0.0),
    (3, 0, ((-1, -1), (1, 1), (-1, 1)), True, 2e3),
]


@pytest.mark.parametrize("case", CASES)
def test_vcycle_matches_transliteration_bitwise(case):
    p = _tiny(*case)
    ref = Ref(p)
    o = Oracle.for_problem(p)
    fields = []
    for impl in ("ref", "oracle"):
        u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
        rng = np.random.default_rng(3)
        for layer in p.tree:
            for _, blk in layer.pairs:
                u.valid(blk)[...] = rng.standard_normal((p.M,) * p.D) * 1e-3
        for l in range(1, len(p.tree) + 1):
            (ref.fill_ghost if impl == "ref" else o.fill_ghost)(u, l)
        for _ in range(2):
            if impl == "ref":
                ref.vcycle(u, b, r, u1, p.s, ndown=2, ntop=3, nup=1, omega=1.0)
            else:
                o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=3, nup=1, omega=1.0)
        fields.append((u, b))
    (ru, rb), (ou, ob) = fields
    G, M = p.b.G, p.M
    vs = (slice(G, G + M),) * p.D
    np.testing.assert_array_equal(ru.u[vs], ou.u[vs])
    np.testing.assert_array_equal(rb.u[vs], ob.u[vs])
    # ghost layers as the reference leaves them
    np.testing.assert_array_equal(ru.u, ou.u)


@pytest.mark.parametrize("case", CASES[:1] + CASES[3:4])
def test_fmg_matches_transliteration_bitwise(case):
    p = _tiny(*case)
    ref = Ref(p)
    o = Oracle.for_problem(p)
    out = []
    for impl in ("ref", "oracle"):
        u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
        for l in range(1, len(p.tree) + 1):
            (ref.fill_ghost if impl == "ref" else o.fill_ghost)(u, l)
        if impl == "ref":
            ref.fmg(u, b, r, u1, p.s, ndown=1, ntop=2, nup=1)
        else:
            o.fmg(u, b, r, u1, p.s, ndown=1, ntop=2, nup=1)
        out.append(u)
    np.testing.assert_array_equal(out[0].u, out[1].u)


def _tiny4(D, geom, bcd):
    """Tiny trees properly nested for the BoxStencil: 2D three levels (a 2x2
    patch of level-1 blocks refined, then one interior level-2 block), 3D two
    levels (one refined block)."""
    if D == 2:
        rootsize, M, levels = (4, 4), 4, 3

        def crit(l, c):
            if l == 1:
                return np.all((c >= 2) & (c <= 3), axis=1)
            return np.all(c == np.array([4, 4]), axis=1)
    else:
        rootsize, M, levels = (2, 2, 2), 4, 2

        def crit(l, c):
            return np.all(c == np.array([1, 2, 1]), axis=1)
    return configs.uniform_problem(D, rootsize, M, levels, geom, bcd, seed=5, refine=crit)


CASES4 = [
    (2, 0, ((1, 1), (-1, -1))),
    (2, 1, ((1, -1), (-1, -1))),
    (2, 2, ((-1, 1), (1, -1))),
    (3, 0, ((1, 1), (1, -1), (-1, -1))),
]


@pytest.mark.parametrize("case", CASES4)
def test_order4_vcycle_matches_transliteration_bitwise(case):
    """The four colours in the reference's order with the 3D z-shift of their
    parities, the compact stencils and the BoxStencil ghost fills: two
    V-cycles of the transliteration and of the oracle, bit for bit."""
    from srmg.tree import box_stencil, connectivity
    p = _tiny4(*case)
    conn = connectivity(p.tree, box_stencil(p.D))
    ref = Ref4(p, conn)
    o = Oracle(p.tree, conn, p.M, p.b.G, p.geom, p.bc, order=4)
    fields = []
    for impl in ("ref", "oracle"):
        u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
        rng = np.random.default_rng(8)
        for layer in p.tree:
            for _, blk in layer.pairs:
                u.valid(blk)[...] = rng.standard_normal((p.M,) * p.D) * 1e-3
        for l in range(1, len(p.tree) + 1):
            (ref.fill_ghost if impl == "ref" else o.fill_ghost)(u, l)
        for _ in range(2):
            if impl == "ref":
                ref.vcycle(u, b, r, u1, p.s, ndown=2, ntop=2, nup=1, omega=1.0)
            else:
                o.vcycle(u, b, r, u1, p.s, ndown=2, ntop=2, nup=1, omega=1.0)
        fields.append((u, b))
    (ru, rb), (ou, ob) = fields
    np.testing.assert_array_equal(ru.u, ou.u)
    np.testing.assert_array_equal(rb.u, ob.u)


def test_order4_colour_order_and_z_shift_are_observable():
    """The pin above can see the colour order and the 3D parity shift: with the
    passes reversed, or without the z shift, the transliteration's 3D result
    differs from the oracle's."""
    from srmg.tree import box_stencil, connectivity
    p = _tiny4(3, 0, ((1, 1), (1, -1), (-1, -1)))
    conn = connectivity(p.tree, box_stencil(p.D))
    o = Oracle(p.tree, conn, p.M, p.b.G, p.geom, p.bc, order=4)

    class Reversed(Ref4):
        def gauss_seidel_iter(self, u, b, omega, s, n, l):
            for _ in range(n):
                for P1, P2 in ((1, 1), (1, 0), (0, 1), (0, 0)):
                    for c, blk in self.tree[l - 1].pairs:
                        self.gauss_seidel4(u, b, omega, s, c, blk, P1, P2)
                    self.fill_ghost_copy(u, l)
                    self.fill_ghost_bnd(u, l)

    class NoShift(Ref4):
        def gauss_seidel4(self, u, b, omega, s, blkpos, blk, P1, P2):
            G, M = self.G, self.M
            for z in range(G + 1, G + M + 1):
                for i1 in range(G + 1, G + M + 1, 2):
                    for i0 in range(G + 1, G + M + 1, 2):
                        I = (i0 + P1, i1 + P2, z)  # parities without the z term
                        J = self.J(I, blkpos)
                        Lu = self.applystencil(u, blk, I, J, s)
                        self.set(u, I, blk, self.get(u, I, blk)
                                 - omega * (Lu + s * self.get(b, I, blk)) / self.diagelm4(J))

    def run(impl):
        u, r, u1, b = p.b.like(), p.b.like(), p.b.like(), p.b.copy()
        for l in range(1, len(p.tree) + 1):
            o.fill_ghost(u, l)
        impl.vcycle(u, b, r, u1, p.s, ndown=1, ntop=1, nup=1, omega=1.0)
        return u.u

    want = run(o)
    np.testing.assert_array_equal(run(Ref4(p, conn)), want)
    assert not np.array_equal(run(Reversed(p, conn)), want)
    assert not np.array_equal(run(NoShift(p, conn)), want)
