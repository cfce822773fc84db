"""Independent assembly of the order-2 operators on a uniform grid.

Built from the finite-volume definition, not from the oracle or the library:
on a cylindrical radial axis the flux through the cell faces at radii
(J-1)h and Jh is weighted by face radius over centre radius (J-1/2)h; a ghost
behind a domain face equals +1 (Neumann) or -1 (Dirichlet) times the adjacent
cell. The Helmholtz operator subtracts s k^2 on the diagonal.
"""

import numpy as np
import scipy.sparse as sp


def axis_matrix(n, radial, alow, ahigh):
    rows, cols, vals = [], [], []
    for i in range(n):
        J = i + 1
        if radial:
            rc = J - 0.5
            wm, wp = (J - 1) / rc, J / rc
        else:
            wm, wp = 1.0, 1.0
        diag = -(wm + wp)
        if i > 0:
            rows.append(i); cols.append(i - 1); vals.append(wm)
        else:
            diag += wm * alow
        if i < n - 1:
            rows.append(i); cols.append(i + 1); vals.append(wp)
        else:
            diag += wp * ahigh
        rows.append(i); cols.append(i); vals.append(diag)
    return sp.csr_matrix((vals, (rows, cols)), shape=(n, n))


def bc_sign(bc, D, d, side):
    idx, m = 0, 1
    for e in range(D):
        f = (1 if side else -1) if e == d else 0
        idx += (f + 1) * m
        m *= 3
    return float(bc[idx])


def grid_operator(N, geom, bc, helm_shift=0.0):
    """Kronecker-sum operator on a grid of extents N (dimension 1 fastest)."""
    D = len(N)
    rdim = {0: -1, 1: 0, 2: 1}[geom]
    A = None
    for d in range(D):
        T = axis_matrix(N[d], d == rdim, bc_sign(bc, D, d, 0), bc_sign(bc, D, d, 1))
        # column-major flattening: dimension 0 fastest -> kron(..., T_0) last
        parts = [sp.identity(N[e], format="csr") for e in range(D)]
        parts[d] = T
        K = parts[D - 1]
        for e in range(D - 2, -1, -1):
            K = sp.kron(K, parts[e], format="csr")
        A = K if A is None else A + K
    if helm_shift:
        A = A + helm_shift * sp.identity(A.shape[0], format="csr")
    return A.tocsc()


def level_to_grid(field, tree, level, M):
    """Valid cells of a complete level as a column-major grid vector."""
    D = tree.D
    layer = tree[level - 1]
    N = [M * (hi - lo + 1) for lo, hi in layer.domain]
    g = np.zeros(N, order="F")
    G = field.G
    for c, blk in layer.pairs:
        sl = tuple(slice(M * (c[d] - 1), M * c[d]) for d in range(D))
        g[sl] = field.valid(blk)
    return g.ravel(order="F"), N


def grid_to_level(vec, N, field, tree, level, M):
    D = tree.D
    g = vec.reshape(N, order="F")
    for c, blk in tree[level - 1].pairs:
        sl = tuple(slice(M * (c[d] - 1), M * c[d]) for d in range(D))
        field.valid(blk)[...] = g[sl]
