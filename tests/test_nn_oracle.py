"""CPU checks of the network oracle (oracle/nn.py): brute-force convolution,
finite-difference gradients, transpose of the interpolation."""
import itertools

import numpy as np
import pytest

from oracle import nn


def brute_conv(a, w, b, stride):
    cin, nz, nx, ny = a.shape
    oz, ox, oy = nn.out_shape((nz, nx, ny), stride)
    y = np.zeros((w.shape[0], oz, ox, oy))
    for co, z, x, yy in itertools.product(range(w.shape[0]), range(oz), range(ox), range(oy)):
        s = b[co]
        for t, (dz, dx, dy) in enumerate(nn.OFFS):
            Z, X, Y = stride * z + dz, stride * x + dx, stride * yy + dy
            if 0 <= Z < nz and 0 <= X < nx and 0 <= Y < ny:
                s += w[co, :, t] @ a[:, Z, X, Y]
        y[co, z, x, yy] = s
    return y


@pytest.mark.parametrize("stride,shape", [(1, (3, 4, 5)), (2, (5, 3, 7))])
def test_conv_matches_brute_force(stride, shape):
    rng = np.random.default_rng(0)
    a = rng.standard_normal((2,) + shape)
    w = rng.standard_normal((3, 2, 27))
    b = rng.standard_normal(3)
    assert np.allclose(nn.conv3d(a, w, b, stride), brute_conv(a, w, b, stride), atol=1e-12)


@pytest.mark.parametrize("stride,relu", [(1, False), (2, False), (1, True)])
def test_conv_backward_finite_differences(stride, relu):
    rng = np.random.default_rng(1)
    a = rng.standard_normal((2, 5, 3, 5))
    w = rng.standard_normal((3, 2, 27)) * 0.3
    b = rng.standard_normal(3) * 0.1
    gy = rng.standard_normal(nn.conv3d(a, w, b, stride).shape)

    def f(a_, w_, b_):
        return float(np.sum(gy * nn.conv3d(a_, w_, b_, stride, relu)))

    y = nn.conv3d(a, w, b, stride, relu)
    ga, gw, gb = nn.conv3d_backward(a, w, gy, stride, y if relu else None)
    h = 1e-6
    for arr, g in ((a, ga), (w, gw), (b, gb)):
        flat = arr.reshape(-1)
        for i in rng.choice(flat.size, min(6, flat.size), replace=False):
            old = flat[i]
            flat[i] = old + h
            fp = f(a, w, b)
            flat[i] = old - h
            fm = f(a, w, b)
            flat[i] = old
            assert (fp - fm) / (2 * h) == pytest.approx(g.reshape(-1)[i], rel=1e-5, abs=1e-7)


def test_upsample_transpose():
    rng = np.random.default_rng(2)
    a = rng.standard_normal((2, 3, 4, 5))
    g = rng.standard_normal(nn.upsample(a).shape)
    assert np.sum(nn.upsample(a) * g) == pytest.approx(np.sum(a * nn.upsample_t(g)), rel=1e-12)
    # coincident nodes are injected, midpoints averaged
    u = nn.upsample(a)
    assert np.array_equal(u[:, ::2, ::2, ::2], a)
    assert np.allclose(u[:, 1, 0, 0], 0.5 * (a[:, 0, 0, 0] + a[:, 1, 0, 0]))


def test_adam_first_step_is_signed_lr():
    w, g = np.array([1.0, -2.0]), np.array([0.5, -3.0])
    w1, m, v = nn.adam(w, g, np.zeros(2), np.zeros(2), 1, 0.01)
    assert np.allclose(w1, w - 0.01 * np.sign(g), atol=1e-7)


def test_net_backward_finite_differences():
    """The whole oracle network's parameter gradient against central
    differences (the skip connections add gradients of two consumers)."""
    rng = np.random.default_rng(3)
    n = sum(co * ci * 27 + co for _, ci, co, _, _ in nn.LAYERS)
    params = rng.standard_normal(n) * 0.15
    X0 = rng.standard_normal((6, 5, 5, 9))
    Y, A = nn.net_forward(params, X0)
    gY = rng.standard_normal(Y.shape)
    g = nn.net_backward(params, A, gY)
    h = 1e-6
    for i in rng.choice(n, 12, replace=False):
        p1, p2 = params.copy(), params.copy()
        p1[i] += h
        p2[i] -= h
        fd = (np.sum(gY * nn.net_forward(p1, X0)[0]) - np.sum(gY * nn.net_forward(p2, X0)[0])) / (2 * h)
        assert fd == pytest.approx(g[i], rel=1e-4, abs=1e-6)
