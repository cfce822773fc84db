"""Pins of the CPU network restatement (oracle/unet_ref.py), the checker of the
learned first iterate's BF16 kernels: finite differences of the loss, closed
forms of the layers, adjointness of the backward convolution."""

import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "oracle"))
import unet_ref as U  # noqa: E402


@pytest.mark.parametrize("D,grid,widths", [(2, (8, 16), (8, 16, 8)), (3, (4, 8, 8), (8, 8))])
def test_loss_gradient_matches_finite_differences(D, grid, widths):
    rng = np.random.default_rng(3)
    p = U.he_params(D, widths, rng)
    f = U.smooth_field(rng, (2,) + grid)
    u = U.smooth_field(rng, (2,) + grid) * 50.0
    c_out = 30.0
    loss, g = U.loss_and_grad(p, f, u, c_out, D, widths)
    assert g.shape == p.shape and np.isfinite(loss)
    # every convolution: some weights and its biases
    specs = U.layer_specs(widths)
    ntap = 3 ** D
    idx, o = [], 0
    for _, ci, co, _ in specs:
        nw = co * ntap * ci
        idx += list(rng.choice(nw, size=3, replace=False) + o)
        idx += list(o + nw + rng.choice(co, size=1))
        o += nw + co
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (U.loss_and_grad(pp, f, u, c_out, D, widths)[0] - U.loss_and_grad(pm, f, u, c_out, D, widths)[0]) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(1e-3, abs(fd), abs(g[i])) + 1e-10, (i, fd, g[i])


def test_convolution_of_an_impulse_is_the_mirrored_kernel():
    rng = np.random.default_rng(0)
    for D in (2, 3):
        n = 5
        x = np.zeros((1,) + (n,) * D + (3,))
        centre = (0,) + (2,) * D
        x[centre + (1,)] = 1.0
        W = rng.normal(size=(4, 3 ** D, 3))
        out = U.conv(x, W, np.zeros(4))
        for t, off in enumerate(U.taps(D)):
            # out[p] = sum_t W[:, t] x[p + off_t]  =>  the impulse at c shows W[:, t] at p = c - off_t
            p = (0,) + tuple(2 - off[d] for d in reversed(range(D)))
            np.testing.assert_allclose(out[p], W[:, t, 1], rtol=0, atol=1e-15)


def test_pool_of_upsample_is_identity_and_adjoint_pairs():
    rng = np.random.default_rng(1)
    for D in (2, 3):
        x = rng.normal(size=(2,) + (4,) * D + (8,))
        np.testing.assert_allclose(U.pool(U.upsample(x)), x, atol=1e-15)
        y = rng.normal(size=(2,) + (8,) * D + (8,))
        # <up x, y> = <x, up^T y>
        assert np.isclose(np.sum(U.upsample(x) * y), np.sum(x * U.upsample_backward(y)))
        # <pool y, x> = <y, pool^T x>
        assert np.isclose(np.sum(U.pool(y) * x), np.sum(y * U.pool_backward(x, y)))


def test_conv_data_gradient_is_the_adjoint():
    rng = np.random.default_rng(2)
    for D in (2, 3):
        x = rng.normal(size=(2,) + (6,) * D + (8,))
        W = rng.normal(size=(5, 3 ** D, 8))
        z = rng.normal(size=(2,) + (6,) * D + (5,))
        dx, dW, db = U.conv_backward(x, W, z)
        assert np.isclose(np.sum(U.conv(x, W, np.zeros(5)) * z), np.sum(x * dx))
        assert np.isclose(np.sum(U.conv(x, W, np.zeros(5)) * z), np.sum(W * dW))
        np.testing.assert_allclose(db, z.sum(axis=tuple(range(D + 1))))


def test_features_layout():
    f = np.zeros((1, 4, 8))
    f[0, 1, 3] = -2.0
    x, scale = U.features(f)
    assert scale[0] == 2.0 and x[0, 1, 3, 0] == -1.0
    np.testing.assert_allclose(x[0, 0, :, 1], 2 * (np.arange(8) + 0.5) / 8 - 1)   # x coordinate (fastest axis)
    np.testing.assert_allclose(x[0, :, 0, 2], 2 * (np.arange(4) + 0.5) / 4 - 1)   # y coordinate
    assert np.all(x[..., 3:] == 0)


def test_param_count_of_the_cfg4_network():
    # 4 levels, widths 8..64, 3D: about 3.3e5 parameters
    assert U.param_count(3, (8, 16, 32, 64)) == 330264 + sum(co for _, _, co, _ in U.layer_specs((8, 16, 32, 64)))


def test_bf16_rounding_known_values():
    vals = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7 + 2 ** -8, -3.0e-3, 65504.0, 0.0])
    # halfway cases round to the even neighbour
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -0.0030059814453125, 65536.0, 0.0])
    np.testing.assert_array_equal(U.to_bf16(vals), want)


def test_bf16_emulation_stays_close_to_exact():
    rng = np.random.default_rng(5)
    widths = (8, 16)
    p = U.he_params(2, widths, rng)
    f = U.smooth_field(rng, (2, 8, 16))
    u = U.smooth_field(rng, (2, 8, 16)) * 10
    l0, g0 = U.loss_and_grad(p, f, u, 10.0, 2, widths)
    l1, g1 = U.loss_and_grad(p, f, u, 10.0, 2, widths, bf16=True)
    assert abs(l1 - l0) < 1e-2 * l0 and l1 != l0
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)
