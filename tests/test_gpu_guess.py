"""The learned first iterate on the device: BF16 matrix-core network kernels
against the CPU restatement (oracle/unet_ref.py), the exact level-1 grid solve
against a sparse direct solve, first iterates inside the solvers (gain over
the zero iterate, fallback to zero on a bad first iterate).

Tolerances: the network stores weights, activations and gradients in BF16
(8-bit mantissa) and accumulates in FP32. Against the restatement that rounds
to BF16 at the same points (unet_ref bf16=True) only the accumulation order
differs, which moves a few values to the neighbouring BF16 number: outputs
agree to 2e-3 in the 2-norm (1e-2 max-norm), loss to 2e-3, gradients to 2e-2
per convolution. Against the exact FP64 restatement: outputs 3e-2 max-norm,
loss 2e-2, gradients 1e-1 (weight gradients sum many cancelling terms). The
FP64 grid solve: 1e-10 relative."""

import os
import sys

import numpy as np
import pytest
import scipy.sparse.linalg as spla

from srmg import configs, solver
from srmg.solver import F32, F64

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
sys.path.insert(0, os.path.dirname(__file__))
import unet_ref as U  # noqa: E402
from assemble import grid_operator  # noqa: E402

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

CASES = [(2, (32, 16), (8, 16, 32)), (3, (16, 16, 8), (8, 16, 32)), (3, (16, 16, 16), (8, 16, 32, 64))]


def _dev(a):
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float32, device="cuda")


def _grid_shape(n):
    return tuple(reversed(n))  # numpy C order: x (n[0]) fastest


@pytest.mark.parametrize("D,n,widths", CASES)
def test_forward_matches_cpu_restatement(D, n, widths):
    print(f"forward D={D} n={n} widths={widths}", flush=True)
    net = solver.UNet(D, n, len(widths), widths, seed=5)
    rng = np.random.default_rng(7)
    p = U.he_params(D, widths, rng, np.float32)
    net.set_params(p)
    net.output_scale = 3.0
    B = 3
    f = U.smooth_field(rng, (B,) + _grid_shape(n))
    fd = _dev(f)
    ud = torch.zeros_like(fd)
    net.forward(fd, ud, B)
    got = ud.cpu().numpy().astype(np.float64)
    errs = {}
    for bf16 in (True, False):
        y, ca = U.forward(p.astype(np.float64), f, D, widths, bf16=bf16)
        ref = y * ca["scale"].reshape((B,) + (1,) * D) * 3.0
        emax = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        el2 = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        errs[bf16] = (emax, el2)
        print(f"  reference rounds to BF16: {bf16}: relative error max-norm {emax:.3e}, 2-norm {el2:.3e}")
    assert errs[True][1] < 2e-3 and errs[True][0] < 1e-2
    assert errs[False][0] < 3e-2


@pytest.mark.parametrize("D,n,widths", CASES)
def test_loss_and_gradient_match_cpu_restatement(D, n, widths):
    print(f"gradient D={D} n={n} widths={widths}", flush=True)
    net = solver.UNet(D, n, len(widths), widths, seed=5)
    rng = np.random.default_rng(11)
    p = U.he_params(D, widths, rng, np.float32)
    net.set_params(p)
    net.output_scale = 20.0
    B = 2
    f = U.smooth_field(rng, (B,) + _grid_shape(n))
    u = U.smooth_field(rng, (B,) + _grid_shape(n)) * 30.0
    loss, g = net.loss_grad(_dev(f), _dev(u), B)
    worst = {}
    for bf16 in (True, False):
        rloss, rg = U.loss_and_grad(p.astype(np.float64), f, u, 20.0, D, widths, bf16=bf16)
        print(f"  reference rounds to BF16: {bf16}; loss {loss:.6e} vs {rloss:.6e}")
        worst[bf16] = [abs(loss - rloss) / rloss, 0.0]
        o = 0
        ntap = 3 ** D
        for k, (_, ci, co, _) in enumerate(U.layer_specs(widths)):
            m = co * ntap * ci + co
            a, b = g[o:o + m].astype(np.float64), rg[o:o + m]
            rel = np.linalg.norm(a - b) / np.linalg.norm(b)
            print(f"  conv {k:2d} ({ci:2d}->{co:2d}) gradient relative error {rel:.3e}")
            worst[bf16][1] = max(worst[bf16][1], rel)
            o += m
    assert worst[True][0] < 2e-3 and worst[True][1] < 2e-2
    assert worst[False][0] < 2e-2 and worst[False][1] < 1e-1


def test_training_reduces_the_loss():
    D, n, widths = 2, (32, 16), (8, 16, 32)
    net = solver.UNet(D, n, 3, widths, seed=1)
    rng = np.random.default_rng(0)
    B = 8
    f = U.smooth_field(rng, (B,) + _grid_shape(n))
    u = np.cumsum(f, axis=-1) * 0.5 + 3.0 * f   # some smooth target map
    net.output_scale = float(np.max(np.abs(u)) / np.max(np.abs(f)))
    fd, ud = _dev(f), _dev(u)
    first = net.train_step(fd, ud, B, 2e-3, want_loss=True)
    last = first
    for _ in range(150):
        last = net.train_step(fd, ud, B, 2e-3, want_loss=True)
    print(f"loss {first:.4e} -> {last:.4e}")
    assert last < 0.2 * first


@pytest.mark.parametrize("D,n,geom,bcdesc,helm", [
    (3, (16, 8, 8), solver.CartesianGeometry(), ((1, 1), (1, 1), (-1, -1)), 0.0),
    (2, (16, 32), solver.CylindricalGeometry(1), ((1, -1), (-1, -1)), 0.0),
    (2, (16, 32), solver.CylindricalGeometry(1), ((1, 1), (-1, -1)), 3.0e5),
])
def test_grid_solve_matches_sparse_direct_solve(D, n, geom, bcdesc, helm):
    from srmg import tree as T
    bc = T.boundaryconditions(bcdesc)
    p = configs.uniform_problem(D, tuple(x // 8 for x in n), 8, 1, geom.code, bcdesc)
    ctx = solver.Context(p.tree, p.conn, p.M)
    rng = np.random.default_rng(4)
    ng = 3
    f = rng.normal(size=(ng, int(np.prod(n))))
    s = 2.5e-6
    disc = solver.HelmholtzDiscretization(D, 2, helm) if helm else solver.LaplacianDiscretization(D, 2)
    fd = torch.tensor(f, dtype=torch.float64, device="cuda")
    ud = torch.zeros_like(fd)
    solver.grid_solve(ctx, D, n, s, geom, bc, disc, fd, ud, ng)
    torch.cuda.synchronize()
    A = grid_operator(n, geom.code, bc, -s * helm if helm else 0.0)
    for g in range(ng):
        ref = spla.spsolve(A, f[g])
        got = ud[g].cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-10 * np.max(np.abs(ref))


def _small_cfg4_like(levels=3):
    """A 3D Cartesian tree with the cfg4 level-1 grid (64^3 in 8^3 blocks) and
    two refined levels around seeded filaments."""
    return configs.make_cfg4(0, scale=0.35, levels=levels)


def _solve(p, mode, net=None, mixed=False, fallback_scale=None):
    ctx = solver.Context(p.tree, p.conn, p.M)
    geom = solver.geometry_from_code(p.geom)
    disc = solver.LaplacianDiscretization(p.D, 2)
    if net is not None:
        solver.set_guess(ctx, net)
    u, b = ctx.field(F64), ctx.field(F64)
    b.upload(p.b)
    if mixed:
        e, be, re, ue = (ctx.field(F32) for _ in range(4))
        n, hist = solver.solve_mixed(u, b, e, be, re, ue, p.s, geom, p.bc, disc, tol=1e-8, guess=mode)
    else:
        r, u1 = ctx.field(F64), ctx.field(F64)
        n, hist = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-8, guess=mode)
    out = p.b.like()
    u.download(out)
    return n, hist, out.u, solver.guess_fallbacks(ctx)


def test_coarse_first_iterate_lowers_the_residual_and_keeps_the_solution():
    # with the exact level-1 solve inside every V-cycle a level-1 first
    # iterate gains less than a cycle at 1e-8: the residual norm is dominated
    # by the finer levels' charge, which no level-1 iterate balances
    p = _small_cfg4_like()
    print(f"tree {[len(l) for l in p.tree]} blocks", flush=True)
    n0, h0, u0, _ = _solve(p, solver.GUESS_NONE)
    n1, h1, u1, fb = _solve(p, solver.GUESS_COARSE)
    print(f"cycles: zero {n0}, coarse {n1}; first residual {h1[0]:.3e}; "
          f"after one cycle {h0[1]:.3e} vs {h1[1]:.3e}")
    assert fb == 0 and h1[0] < 1.0 and n1 <= n0 and h1[-1] <= 1e-8
    assert np.all(h1[1:n1 + 1] < h0[1:n1 + 1])
    G, M = p.b.G, p.M
    v = (slice(G, G + M),) * 3
    err = np.max(np.abs(u1[v] - u0[v])) / np.max(np.abs(u0[v]))
    assert err < 1e-6


def test_bad_learned_first_iterate_falls_back_to_zero():
    p = _small_cfg4_like()
    net = solver.UNet(3, (64, 64, 64), 4, (8, 16, 32, 64), seed=3)
    net.set_problem(solver.geometry_from_code(p.geom), p.bc, solver.LaplacianDiscretization(3, 2))
    net.output_scale = 1e9   # a first iterate far larger than the solution
    n0, h0, u0, _ = _solve(p, solver.GUESS_NONE)
    n1, h1, u1, fb = _solve(p, solver.GUESS_LEARNED, net)
    assert fb == 1
    np.testing.assert_array_equal(h1, h0)
    np.testing.assert_array_equal(u1, u0)


def test_learned_first_iterate_refuses_another_operator():
    p = _small_cfg4_like()
    net = solver.UNet(3, (64, 64, 64), 4, (8, 16, 32, 64), seed=3)
    from srmg import tree as T
    other = T.boundaryconditions(((1, 1), (1, 1), (1, 1)))
    net.set_problem(solver.geometry_from_code(p.geom), other, solver.LaplacianDiscretization(3, 2))
    with pytest.raises(solver.SrmgError, match="another level-1 operator"):
        _solve(p, solver.GUESS_LEARNED, net)
    with pytest.raises(solver.SrmgError, match="no network"):
        _solve(p, solver.GUESS_LEARNED, None)


def _trained():
    from srmg import guess
    path = guess.cfg4_weights_path()
    if not os.path.exists(path):
        pytest.skip("no trained weights in the package")
    return guess.load(path)


def test_trained_network_predicts_held_out_level1_solutions():
    from srmg import guess
    net = _trained()
    gp = guess.cfg4_grid_problem()
    small = configs.uniform_problem(3, (1, 1, 1), 8, 1, 0, ((1, 1), (1, 1), (-1, -1)))
    ctx = solver.Context(small.tree, small.conn, small.M)
    F, U = guess.make_dataset(ctx, gp, 32, seed0=10_000_000, log=lambda m: None)
    err = guess.evaluate(net, F, U)
    print(f"held-out relative error: median {np.median(err):.3f}, max {err.max():.3f}")
    assert np.median(err) < 0.2


def test_learned_first_iterate_cuts_cycles_with_the_reference_coarse_treatment():
    # the reference's level-1 treatment (ntop Gauss-Seidel sweeps) resolves the
    # 64^3 level 1 slowly; a learned level-1 first iterate reaches the residual
    # of 100 zero-start cycles in a quarter of the cycles
    net = _trained()
    p = _small_cfg4_like()
    geom, disc = solver.geometry_from_code(p.geom), solver.LaplacianDiscretization(3, 2)
    res = {}
    for mode in (solver.GUESS_NONE, solver.GUESS_LEARNED):
        ctx = solver.Context(p.tree, p.conn, p.M)
        solver.set_guess(ctx, net)
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        b.upload(p.b)
        n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=1e-12, max_cycles=100,
                            coarse=solver.COARSE_GS, guess=mode)
        res[mode] = (h, solver.guess_fallbacks(ctx))
    h0, (h1, fb) = res[solver.GUESS_NONE][0], res[solver.GUESS_LEARNED]
    k = int(np.argmax(h1 <= h0[100]))
    print(f"zero start: {h0[100]:.3e} after 100 cycles; learned start reaches it after {k} cycles "
          f"({h1[100]:.3e} after 100)")
    assert fb == 0 and h1[k] <= h0[100] and k <= 25


def test_learned_first_iterate_with_the_exact_level1_solve_converges_to_the_same_solution():
    net = _trained()
    p = _small_cfg4_like()
    n0, h0, u0, _ = _solve(p, solver.GUESS_NONE)
    n1, h1, u1, fb = _solve(p, solver.GUESS_LEARNED, net)
    print(f"cycles: zero {n0}, learned {n1}; first residual {h1[0]:.3e}")
    assert fb == 0 and h1[0] < 1.0 and n1 <= n0 and h1[-1] <= 1e-8
    G, M = p.b.G, p.M
    v = (slice(G, G + M),) * 3
    assert np.max(np.abs(u1[v] - u0[v])) / np.max(np.abs(u0[v])) < 1e-6


# ---- the 2D network of the cylindrical streamer workload (cfg2's level 1) ----

def _trained2d():
    from srmg import guess
    path = guess.weights_path("cfg2")
    if not os.path.exists(path):
        pytest.skip("no trained 2D weights in the package")
    return guess.load(path)


def test_trained_2d_network_predicts_held_out_level1_solutions():
    from srmg import guess
    net = _trained2d()
    gp = guess.cfg2_grid_problem()
    assert net.D == 2 and tuple(net.n) == gp.n
    small = configs.uniform_problem(2, (1, 1), 8, 1, 0, ((1, 1), (1, 1)))
    ctx = solver.Context(small.tree, small.conn, small.M)
    F, U = guess.make_dataset(ctx, gp, 64, seed0=10_000_000, log=lambda m: None)
    err = guess.evaluate(net, F, U)
    print(f"2D held-out relative error: median {np.median(err):.3f}, 90th percentile {np.percentile(err, 90):.3f}")
    assert np.median(err) < 0.2


@pytest.mark.parametrize("coarse", [solver.COARSE_EXACT, solver.COARSE_GS])
def test_2d_learned_first_iterate_on_the_streamer_tree(coarse):
    # the full cfg2 tree: the learned start is accepted (residual below the
    # zero start's) and converges to the zero start's solution with the exact
    # level-1 solve; with the reference's level-1 sweeps on the 128 x 256
    # level 1 its first cycles stay below the zero start's
    net = _trained2d()
    p = configs.make_cfg2(0)
    geom, disc = solver.geometry_from_code(p.geom), solver.LaplacianDiscretization(2, 2)
    tol, maxc = (1e-8, 40) if coarse == solver.COARSE_EXACT else (1e-30, 10)
    res = {}
    for mode in (solver.GUESS_NONE, solver.GUESS_LEARNED):
        ctx = solver.Context(p.tree, p.conn, p.M)
        solver.set_guess(ctx, net)
        u, b, r, u1 = (ctx.field(F64) for _ in range(4))
        b.upload(p.b)
        n, h = solver.solve(u, b, r, u1, p.s, geom, p.bc, disc, tol=tol, max_cycles=maxc, coarse=coarse,
                            guess=mode)
        out = p.b.like()
        u.download(out)
        res[mode] = (n, np.asarray(h).copy(), out.u.copy(), solver.guess_fallbacks(ctx))
    (n0, h0, u0, _), (n1, h1, u1, fb) = res[solver.GUESS_NONE], res[solver.GUESS_LEARNED]
    print(f"coarse={coarse}: zero start {n0} cycles {np.array2string(h0, precision=3)}; learned start {n1} "
          f"cycles {np.array2string(h1, precision=3)}, fallbacks {fb}", flush=True)
    assert fb == 0 and h1[0] < 1.0
    if coarse == solver.COARSE_EXACT:
        assert h1[-1] <= 1e-8 and n1 <= n0
        G, M = p.b.G, p.M
        v = (slice(G, G + M),) * 2
        assert np.max(np.abs(u1[v] - u0[v])) / np.max(np.abs(u0[v])) < 1e-6
    else:
        assert np.all(h1[1:4] < h0[1:4])
