"""srmg_grid_solve_noalloc: the exact stacked-grid solve with its operands and
intermediate grids in the library's fixed device scratch array instead of
hipMalloc'd buffers. Its results equal srmg_grid_solve's bit for bit, on
consecutive calls that reuse the array with a smaller problem the second
time, and a problem above the array's size is refused before anything runs.
"""

import numpy as np
import pytest

torch = pytest.importorskip("torch")  # noqa: F841  (torch's HIP runtime first)

from srmg import configs, solver
from srmg import tree as T

pytestmark = pytest.mark.gpu

# (D, n, geometry, boundary table, Helmholtz k2, grids): the second case is
# smaller than the first, so the second call reads a scratch array that still
# holds the first call's larger grids
CASES = [
    (3, (64, 64, 64), solver.CartesianGeometry(), ((1, 1), (1, 1), (-1, -1)), 0.0, 2),
    (2, (16, 32), solver.CylindricalGeometry(1), ((1, 1), (-1, -1)), 3.0e5, 3),
]


def _context(D, n, geom, bcdesc):
    p = configs.uniform_problem(D, tuple(x // 8 for x in n), 8, 1, geom.code, bcdesc)
    return solver.Context(p.tree, p.conn, p.M)


def _solve_both(case, seed):
    D, n, geom, bcdesc, helm, ng = case
    bc = T.boundaryconditions(bcdesc)
    disc = solver.HelmholtzDiscretization(D, 2, helm) if helm else solver.LaplacianDiscretization(D, 2)
    ctx = _context(D, n, geom, bcdesc)
    g = torch.Generator(device="cpu").manual_seed(seed)
    f = torch.randn((ng, int(np.prod(n))), dtype=torch.float64, generator=g).cuda()
    u_ref = torch.full_like(f, float("nan"))
    u_new = torch.full_like(f, float("nan"))
    torch.cuda.synchronize()
    solver.grid_solve(ctx, D, n, 2.5e-6, geom, bc, disc, f, u_ref, ng)
    solver.grid_solve(ctx, D, n, 2.5e-6, geom, bc, disc, f, u_new, ng, noalloc=True)
    torch.cuda.synchronize()
    return u_ref.cpu().numpy(), u_new.cpu().numpy()


def test_noalloc_grid_solve_equals_grid_solve_on_consecutive_calls():
    for k, case in enumerate(CASES):
        print(f"grid_solve_noalloc case {k}: D={case[0]} n={case[1]} grids={case[5]}", flush=True)
        ref, got = _solve_both(case, seed=11 + k)
        assert np.all(np.isfinite(ref))
        np.testing.assert_array_equal(got, ref)


def test_noalloc_grid_solve_refuses_grids_above_the_scratch_size():
    D, n, geom, bcdesc = 3, (64, 64, 64), solver.CartesianGeometry(), ((1, 1), (1, 1), (-1, -1))
    ctx = _context(D, n, geom, bcdesc)
    ng = 4  # two 4 x 64^3 FP64 grids are 16 MiB alone
    f = torch.zeros((ng, int(np.prod(n))), dtype=torch.float64, device="cuda")
    u = torch.full_like(f, 7.0)
    torch.cuda.synchronize()
    with pytest.raises(solver.SrmgError, match="maximum"):
        solver.grid_solve(ctx, D, n, 1.0, geom, T.boundaryconditions(bcdesc), solver.LaplacianDiscretization(D, 2),
                          f, u, ng, noalloc=True)
    torch.cuda.synchronize()
    assert bool(torch.all(u == 7.0))
