"""Training data and training loop of the learned first iterate.

Samples are level-1 problems of one geometry: seeded sums of 1 to 12
Gaussian blobs and filaments (line segments with a Gaussian radial profile),
amplitudes log-uniform in [1e18, 1e21] m^-3, some of negative sign; the target
of each is the exact solution of L_1 u = f, f = -s b, from srmg_grid_solve
(the library's FP64 level-1 solve). Charges are evaluated on the device with
torch tensor arithmetic (input synthesis only); solves, training steps and
inference are the library's kernels.
"""

from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import solver

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_DIR = os.path.join(HERE, "weights")


@dataclass
class GridProblem:
    """The level-1 problem a network is trained for."""
    D: int
    n: tuple          # cells per dimension (x first)
    L: tuple          # domain length per dimension [m]
    geometry: object  # solver.CartesianGeometry / CylindricalGeometry
    bc: np.ndarray    # 3^D table (src/boundary.jl layout)
    s: float          # h^2 e / eps0 of level 1

    @property
    def h(self):
        return self.L[0] / self.n[0]


def cfg4_grid_problem() -> GridProblem:
    """Level 1 of the 3D Cartesian workload: 64^3 cells on 1 cm^3, BC
    ((1,1),(1,1),(-1,-1))."""
    from .configs import ELEMENTARY_CHARGE, EPSILON_0
    from .tree import boundaryconditions
    h = 1e-2 / 64
    return GridProblem(3, (64, 64, 64), (1e-2,) * 3, solver.CartesianGeometry(),
                       boundaryconditions(((1, 1), (1, 1), (-1, -1))), h * h * ELEMENTARY_CHARGE / EPSILON_0)


def cfg2_grid_problem() -> GridProblem:
    """Level 1 of the 2D cylindrical streamer workload: 128 x 256 cells on
    r in [0, 1 cm], z in [0, 2 cm] (CylindricalGeometry{1}, r first), BC
    ((1,-1),(-1,-1)): Neumann on the axis, Dirichlet at the outer radius and
    the electrodes."""
    from .configs import ELEMENTARY_CHARGE, EPSILON_0
    from .tree import boundaryconditions
    h = 1e-2 / 128
    return GridProblem(2, (128, 256), (1e-2, 2e-2), solver.CylindricalGeometry(1),
                       boundaryconditions(((1, -1), (-1, -1))), h * h * ELEMENTARY_CHARGE / EPSILON_0)


def grid_problem(config: str) -> GridProblem:
    return {"cfg2": cfg2_grid_problem, "cfg4": cfg4_grid_problem}[config]()


def synth_charges(gp: GridProblem, seeds: Sequence[int], device="cuda"):
    """Charge densities [len(seeds), V] (float64, x fastest) on the device;
    sample k is fixed by seeds[k]."""
    import torch
    D = gp.D
    axes = [(torch.arange(gp.n[d], device=device, dtype=torch.float64) + 0.5) * (gp.L[d] / gp.n[d])
            for d in range(D)]
    # meshgrid with x fastest: arrays indexed [.., z, y, x]
    mesh = torch.meshgrid(*reversed(axes), indexing="ij")
    X = [mesh[D - 1 - d].reshape(-1) for d in range(D)]
    out = torch.zeros((len(seeds), X[0].numel()), dtype=torch.float64, device=device)
    Lv = np.array(gp.L[:D])
    # cylindrical grids: half of the charges sit on the axis (r < 0.05 L_r;
    # their filaments run along z), as streamer channels and heads do
    rdim = {1: 0, 2: 1}.get(int(gp.geometry.code), -1)
    for k, seed in enumerate(seeds):
        rng = np.random.default_rng(int(seed))
        acc = torch.zeros_like(X[0])
        for _ in range(int(rng.integers(1, 13))):
            amp = 10 ** rng.uniform(18, 21) * (-1 if rng.uniform() < 0.3 else 1)
            axial = rdim >= 0 and rng.uniform() < 0.5
            if rng.uniform() < 0.5:
                c = rng.uniform(0.1, 0.9, D) * Lv
                if axial:
                    c[rdim] = rng.uniform(0.0, 0.05) * Lv[rdim]
                w = rng.uniform(0.2e-3, 1.0e-3)
                r2 = sum((X[d] - c[d]) ** 2 for d in range(D))
                acc += amp * torch.exp(-r2 / (w * w))
            else:
                a = rng.uniform(0.1, 0.9, D) * Lv
                dirn = rng.normal(size=D)
                if axial:
                    a[rdim] = rng.uniform(0.0, 0.05) * Lv[rdim]
                    dirn[rdim] = 0.1 * dirn[rdim]
                dirn /= np.linalg.norm(dirn)
                b = np.clip(a + dirn * rng.uniform(1e-3, 5e-3), 0.02 * Lv, 0.98 * Lv)
                ab = b - a
                ab2 = float(ab @ ab) or 1e-30
                t = sum((X[d] - a[d]) * ab[d] for d in range(D)) / ab2
                t = t.clamp(0.0, 1.0)
                r2 = sum((X[d] - a[d] - t * ab[d]) ** 2 for d in range(D))
                rad = rng.uniform(0.1e-3, 0.3e-3)
                acc += amp * torch.exp(-r2 / (rad * rad))
        out[k] = acc
    return out


def make_dataset(ctx, gp: GridProblem, nsamples: int, seed0: int = 0, chunk: int = 128, device="cuda",
                 log=print):
    """(f, u) float32 device tensors [nsamples, V] for L_1 u = f, f = -s b."""
    import torch
    V = int(np.prod(gp.n))
    F = torch.empty((nsamples, V), dtype=torch.float32, device=device)
    U = torch.empty((nsamples, V), dtype=torch.float32, device=device)
    disc = solver.LaplacianDiscretization(gp.D, 2)
    t0 = time.time()
    for i in range(0, nsamples, chunk):
        k = min(chunk, nsamples - i)
        b = synth_charges(gp, range(seed0 + i, seed0 + i + k), device)
        f = (-gp.s) * b
        u = torch.empty_like(f)
        torch.cuda.synchronize()
        solver.grid_solve(ctx, gp.D, gp.n, gp.s, gp.geometry, gp.bc, disc, f, u, k)
        F[i:i + k] = f.float()
        U[i:i + k] = u.float()
        if (i // chunk) % 20 == 0:
            log(f"[dataset] {i + k}/{nsamples} samples, {time.time() - t0:.1f} s")
    torch.cuda.synchronize()
    return F, U


def output_scale(F, U) -> float:
    """Median of max|u| / max|f| over the samples: the network then predicts O(1) values."""
    r = U.abs().amax(dim=1).double() / F.abs().amax(dim=1).double().clamp_min(1e-300)
    return float(r.median())


def train(net: solver.UNet, F, U, *, epochs=2, batch=16, lr=1e-3, lr_min=5e-5, seed=0, log=print,
          log_every=100):
    """Adam with cosine learning-rate decay over shuffled batches; returns the
    loss history (one value per log interval)."""
    import torch
    n = F.shape[0]
    steps = epochs * (n // batch)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    t0 = time.time()
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(n, generator=g).to(F.device)
        for i in range(0, n - batch + 1, batch):
            idx = perm[i:i + batch]
            fb = F.index_select(0, idx).contiguous()
            ub = U.index_select(0, idx).contiguous()
            torch.cuda.synchronize()
            lr_t = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(1, steps)))
            want = (step % log_every) == 0
            loss = net.train_step(fb, ub, batch, lr_t, want_loss=want)
            if want:
                hist.append((step, loss))
                log(f"[train] epoch {ep} step {step}/{steps} loss {loss:.4e} lr {lr_t:.2e} "
                    f"{time.time() - t0:.1f} s")
            step += 1
    net.synchronize()
    return hist


def evaluate(net: solver.UNet, F, U, batch=16):
    """Relative 2-norm error of the prediction per sample (device tensors in)."""
    import torch
    errs = []
    for i in range(0, F.shape[0], batch):
        k = min(batch, F.shape[0] - i)
        fb = F[i:i + k].contiguous()
        out = torch.empty_like(fb)
        net.forward(fb, out, k)
        ub = U[i:i + k]
        errs.append(((out - ub).norm(dim=1) / ub.norm(dim=1)).double().cpu().numpy())
    return np.concatenate(errs)


def save(net: solver.UNet, gp: GridProblem, path: str, meta: Optional[dict] = None):
    os.makedirs(os.path.dirname(path), exist_ok=True)
    np.savez(path, params=net.get_params(), output_scale=net.output_scale, D=gp.D, n=np.array(gp.n),
             levels=net.levels, widths=np.array(net.widths), geometry=int(gp.geometry.code),
             bc=np.asarray(gp.bc, np.int8).ravel(), **(meta or {}))


def load(path: str) -> solver.UNet:
    """A network from save(); it is declared for its level-1 operator (Laplacian)."""
    z = np.load(path, allow_pickle=False)
    D = int(z["D"])
    net = solver.UNet(D, tuple(int(x) for x in z["n"]), int(z["levels"]), tuple(int(x) for x in z["widths"]))
    net.set_params(z["params"])
    net.output_scale = float(z["output_scale"])
    net.set_problem(solver.geometry_from_code(int(z["geometry"])), z["bc"], solver.LaplacianDiscretization(D, 2))
    return net


def weights_path(config: str) -> str:
    """The committed network of a workload's level-1 grid (cfg2: 2D 128 x 256
    cylindrical, cfg4: 3D 64^3 Cartesian)."""
    return os.path.join(WEIGHTS_DIR, f"unet_{config}_level1.npz")


def cfg4_weights_path() -> str:
    return weights_path("cfg4")
