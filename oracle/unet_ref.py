"""CPU restatement of the learned first iterate's network, for the tests only
(the checker of the HIP kernels; nothing in the library calls it).

The network has no counterpart in the reference: it is textbook arithmetic
written with numpy in any float dtype (float64 for finite-difference checks,
float32 as the reference of the BF16 kernels):

* 3^D convolution with zero padding, weights W[cout][tap][cin], taps
  t = (dx+1) + 3 (dy+1) + 9 (dz+1) (dx along the fastest grid axis), bias;
* ReLU; 2^D average pooling; nearest-neighbour upsampling; additive skip;
* mean squared error against u / (max|f| c_out); gradients by the chain rule.

Grids are numpy arrays [batch, (n2,) n1, n0] (C order, x = n0 fastest), the
same memory order as the device's stacked grids; activations carry a trailing
channel axis.

Parity pins (tests/test_unet_ref.py): finite differences of the loss for every
kind of parameter; the convolution of a unit impulse reproduces the mirrored
weights; pooling of upsampled data is the identity; the data gradient of the
convolution is its adjoint.
"""

from __future__ import annotations

from typing import List, Tuple

import numpy as np

FEAT = 8


def to_bf16(x):
    """Round to the nearest BF16 value (ties to even), via FP32."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return (u.astype(np.uint32)).view(np.float32).astype(np.float64)


def _ident(x):
    return x


def layer_specs(widths) -> List[Tuple[int, int, int, bool]]:
    """(level, cin, cout, relu) of every convolution, in parameter order."""
    L = len(widths)
    specs = []
    for l in range(L):
        specs.append((l, FEAT if l == 0 else widths[l - 1], widths[l], True))
        specs.append((l, widths[l], widths[l], True))
    for l in range(L - 2, -1, -1):
        specs.append((l + 1, widths[l + 1], widths[l], False))
        specs.append((l, widths[l], widths[l], True))
    specs.append((0, widths[0], 1, False))
    return specs


def param_count(D: int, widths) -> int:
    ntap = 3 ** D
    return sum(co * ntap * ci + co for _, ci, co, _ in layer_specs(widths))


def unpack(params, D: int, widths):
    ntap = 3 ** D
    out, o = [], 0
    for _, ci, co, _ in layer_specs(widths):
        W = params[o:o + co * ntap * ci].reshape(co, ntap, ci)
        o += co * ntap * ci
        b = params[o:o + co]
        o += co
        out.append((W, b))
    assert o == len(params)
    return out


def taps(D: int):
    """(dx, dy[, dz]) offsets in tap order (dx fastest)."""
    rng = (-1, 0, 1)
    if D == 2:
        return [(dx, dy) for dy in rng for dx in rng]
    return [(dx, dy, dz) for dz in rng for dy in rng for dx in rng]


def _shift_slices(D, off):
    """Slices of the padded array that hold x[p + off] for every grid point p."""
    # array axes: batch, spatial axes in order (z,) y, x, channel
    sl = [slice(None)]
    for d in reversed(range(D)):  # spatial axis order: highest dim first
        o = off[d]
        sl.append(slice(1 + o, None if o == 1 else o - 1))
    sl.append(slice(None))
    return tuple(sl)


def conv(x, W, b):
    """x [B, *grid, cin] -> [B, *grid, cout]"""
    D = x.ndim - 2
    pad = [(0, 0)] + [(1, 1)] * D + [(0, 0)]
    xp = np.pad(x, pad)
    out = np.zeros(x.shape[:-1] + (W.shape[0],), dtype=x.dtype)
    for t, off in enumerate(taps(D)):
        out += xp[_shift_slices(D, off)] @ W[:, t, :].T
    return out + b


def conv_backward(x, W, dout):
    """Gradients of conv(x, W, b) w.r.t. x, W and b given d out."""
    D = x.ndim - 2
    pad = [(0, 0)] + [(1, 1)] * D + [(0, 0)]
    xp = np.pad(x, pad)
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    for t, off in enumerate(taps(D)):
        sl = _shift_slices(D, off)
        xs = xp[sl]
        dW[:, t, :] = np.tensordot(dout, xs, axes=(list(range(D + 1)), list(range(D + 1))))
        dxp[sl] += dout @ W[:, t, :]
    db = dout.sum(axis=tuple(range(D + 1)))
    dx = dxp[(slice(None),) + (slice(1, -1),) * D + (slice(None),)]
    return dx, dW, db


def pool(x):
    D = x.ndim - 2
    B, C = x.shape[0], x.shape[-1]
    sp = x.shape[1:-1]
    shape = [B]
    for n in sp:
        shape += [n // 2, 2]
    r = x.reshape(shape + [C])
    return r.mean(axis=tuple(2 + 2 * i for i in range(D)))


def upsample(x):
    D = x.ndim - 2
    for a in range(1, D + 1):
        x = np.repeat(x, 2, axis=a)
    return x


def pool_backward(dp, like):
    D = like.ndim - 2
    return upsample(dp) / (2 ** D)


def upsample_backward(dfine):
    D = dfine.ndim - 2
    return pool(dfine) * (2 ** D)


def features(f):
    """f [B, *grid] -> ([B, *grid, 8], scale[B])"""
    D = f.ndim - 1
    B = f.shape[0]
    scale = np.abs(f.reshape(B, -1)).max(axis=1)
    scale = np.where(scale > 0, scale, 1.0)
    x = np.zeros(f.shape + (FEAT,), dtype=f.dtype)
    x[..., 0] = f / scale.reshape((B,) + (1,) * D)
    sp = f.shape[1:]
    for d in range(D):  # coordinate of dim d lives on array axis D - d
        n = sp[D - 1 - d]
        c = 2.0 * (np.arange(n) + 0.5) / n - 1.0
        shape = [1] * (D + 1)
        shape[D - d] = n
        x[..., 1 + d] = np.broadcast_to(c.reshape(shape), f.shape)
    return x, scale


def _weights(params, D, widths, q):
    return [(q(W), b) for W, b in unpack(params, D, widths)]


def forward(params, f, D: int, widths, bf16: bool = False):
    """Normalized prediction y [B, *grid] and the activations for backward.
    bf16: round weights, input features and every stored activation to BF16
    where the device kernels store them (accumulation stays exact here)."""
    q = to_bf16 if bf16 else _ident
    P = _weights(params, D, widths, q)
    L = len(widths)
    x, scale = features(f)
    x = q(x)
    cache = {"x": x}
    A, E, Pp, Cc, S, Dd = {}, {}, {}, {}, {}, {}
    k = 0
    for l in range(L):
        inp = x if l == 0 else Pp.setdefault(l, q(pool(E[l - 1])))
        W, b = P[k]; k += 1
        A[l] = q(np.maximum(conv(inp, W, b), 0))
        W, b = P[k]; k += 1
        E[l] = q(np.maximum(conv(A[l], W, b), 0))
    for l in range(L - 2, -1, -1):
        src = E[L - 1] if l == L - 2 else Dd[l + 1]
        W, b = P[k]; k += 1
        Cc[l] = q(conv(src, W, b))
        S[l] = q(upsample(Cc[l]) + E[l])
        W, b = P[k]; k += 1
        Dd[l] = q(np.maximum(conv(S[l], W, b), 0))
    W, b = P[k]
    y = conv(Dd[0], W, b)[..., 0]
    if bf16:
        y = y.astype(np.float32).astype(np.float64)
    cache.update(A=A, E=E, P=Pp, C=Cc, S=S, Dd=Dd, scale=scale)
    return y, cache


def loss_and_grad(params, f, u, c_out: float, D: int, widths, bf16: bool = False):
    """MSE of the normalized prediction and its gradient (parameter order).
    bf16: also round every stored gradient tensor to BF16 where the device
    kernels store them."""
    q = to_bf16 if bf16 else _ident
    y, ca = forward(params, f, D, widths, bf16)
    B = f.shape[0]
    t = u / (ca["scale"].reshape((B,) + (1,) * D) * c_out)
    d = y - t
    loss = float(np.mean(d * d))
    if bf16:
        dy = q(np.float32(2.0 / d.size) * d.astype(np.float32))
    else:
        dy = (2.0 / d.size) * d
    P = _weights(params, D, widths, q)
    L = len(widths)
    grads = [None] * len(P)
    A, E, Pp, Cc, S, Dd = ca["A"], ca["E"], ca["P"], ca["C"], ca["S"], ca["Dd"]
    # layer indices
    idx_a = [2 * l for l in range(L)]
    idx_e = [2 * l + 1 for l in range(L)]
    idx_c, idx_d = {}, {}
    k = 2 * L
    for l in range(L - 2, -1, -1):
        idx_c[l], idx_d[l] = k, k + 1
        k += 2
    idx_y = k
    gDd, gS, gE = {}, {}, {}
    W, _ = P[idx_y]
    dx, dW, db = conv_backward(Dd[0], W, dy[..., None])
    grads[idx_y] = (dW, db)
    gDd[0] = q(dx * (Dd[0] > 0))
    for l in range(0, L - 1):
        W, _ = P[idx_d[l]]
        dx, dW, db = conv_backward(S[l], W, gDd[l])
        grads[idx_d[l]] = (dW, db)
        gS[l] = q(dx)
        gC = q(upsample_backward(gS[l]))
        src = E[L - 1] if l == L - 2 else Dd[l + 1]
        W, _ = P[idx_c[l]]
        dx, dW, db = conv_backward(src, W, gC)
        grads[idx_c[l]] = (dW, db)
        if l == L - 2:
            gE[L - 1] = q(dx * (src > 0))
        else:
            gDd[l + 1] = q(dx * (src > 0))
    for l in range(L - 1, -1, -1):
        W, _ = P[idx_e[l]]
        dx, dW, db = conv_backward(A[l], W, gE[l])
        grads[idx_e[l]] = (dW, db)
        gA = q(dx * (A[l] > 0))
        inp = ca["x"] if l == 0 else Pp[l]
        W, _ = P[idx_a[l]]
        dx, dW, db = conv_backward(inp, W, gA)
        grads[idx_a[l]] = (dW, db)
        if l > 0:
            gE[l - 1] = q((gS[l - 1] + pool_backward(q(dx), E[l - 1])) * (E[l - 1] > 0))
    flat = np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])
    return loss, flat


def he_params(D: int, widths, rng, dtype=np.float64):
    """He-normal weights and small random biases (test inputs)."""
    ntap = 3 ** D
    out = []
    for _, ci, co, relu in layer_specs(widths):
        sd = np.sqrt((2.0 if relu else 1.0) / (ntap * ci))
        out.append(rng.normal(0, sd, co * ntap * ci))
        out.append(rng.normal(0, 0.05, co))
    return np.concatenate(out).astype(dtype)


def smooth_field(rng, shape, nbump=3):
    """A sum of Gaussian bumps on a [B, *grid] array (test inputs)."""
    B, sp = shape[0], shape[1:]
    grids = np.meshgrid(*[np.arange(n) for n in sp], indexing="ij")
    out = np.zeros(shape)
    for bi in range(B):
        for _ in range(nbump):
            c = [rng.uniform(0, n) for n in sp]
            w = rng.uniform(1.0, max(2.0, min(sp) / 3))
            r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
            out[bi] += rng.uniform(-1, 1) * np.exp(-r2 / w ** 2)
    return out


__all__ = ["layer_specs", "param_count", "unpack", "taps", "conv", "conv_backward", "pool", "upsample",
           "features", "forward", "loss_and_grad", "he_params", "smooth_field", "FEAT"]
