"""Launch the learned first iterate's network kernels for timing and counter
runs: forward passes at batch 1 (inference) and training steps at batch 16 on
a workload's level-1 grid with the committed weights.

    python steamroller.jl_amd/unet_bench.py --config cfg4 --forwards 20 --steps 10

Prints the host-timed milliseconds per forward pass and per training step and
the FLOPs of one forward pass (2 x taps x cin x cout per voxel and
convolution). Under rocprofv3 --kernel-trace the durations of the k_conv and
k_wgrad launches give each kernel's achieved TFLOP/s.
"""

import argparse
import os
import sys
import time

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))


def conv_layers(D, widths):
    """(level, cin, cout) of every convolution in forward order: two per
    encoder level (input features 8 channels), per decoder level a coarse
    convolution and one after the skip, then the output convolution."""
    L = len(widths)
    out = []
    cin = 8
    for l in range(L):
        out += [(l, cin, widths[l]), (l, widths[l], widths[l])]
        cin = widths[l]
    for l in range(L - 2, -1, -1):
        out += [(l + 1, widths[l + 1], widths[l]), (l, widths[l], widths[l])]
    out.append((0, widths[0], 1))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="cfg4", choices=["cfg2", "cfg4"])
    ap.add_argument("--forwards", type=int, default=20)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=16)
    a = ap.parse_args()
    from __graft_entry__ import load_package
    load_package()
    import numpy as np
    import torch
    from srmg import guess
    net = guess.load(guess.weights_path(a.config))
    gp = guess.grid_problem(a.config)
    V = int(np.prod(gp.n))
    taps = 3 ** gp.D
    flops = sum(2 * taps * ci * co * (V >> (gp.D * l)) for l, ci, co in conv_layers(gp.D, net.widths))
    print(f"{a.config}: {gp.D}D grid {gp.n}, widths {net.widths}, {net.nparams} parameters, "
          f"{flops / 1e9:.3f} GFLOP per forward pass at batch 1", flush=True)
    f = guess.synth_charges(gp, range(a.batch)).float().mul_(-gp.s).contiguous()
    y = torch.empty_like(f)
    net.forward(f[:1], y[:1], 1)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(a.forwards):
        net.forward(f[:1], y[:1], 1)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / max(1, a.forwards)
    print(f"forward, batch 1: {ms:.3f} ms ({flops / ms / 1e9:.2f} TFLOP/s host-timed)", flush=True)
    u = (y * 1.0).contiguous()
    net.train_step(f, u, a.batch, 1e-5)
    net.synchronize()
    t = time.perf_counter()
    for _ in range(a.steps):
        net.train_step(f, u, a.batch, 1e-5)
    net.synchronize()
    ms = 1e3 * (time.perf_counter() - t) / max(1, a.steps)
    print(f"training step, batch {a.batch}: {ms:.3f} ms (forward + backward about 3x the forward FLOPs: "
          f"{3 * a.batch * flops / ms / 1e9:.2f} TFLOP/s host-timed)", flush=True)


if __name__ == "__main__":
    main()
