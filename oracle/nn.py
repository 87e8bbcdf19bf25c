"""CPU oracle of the learned coarse-correction network (test infrastructure).

The network has no counterpart in topoptlab; this file restates, in plain
numpy float64, the operations the HIP kernels implement so that their
forward values and gradients can be checked:

* ``conv3d``: 3x3x3 convolution on a node grid, zero padding, stride 1 or 2
  (stride 2 maps 2m+1 nodes to m+1, node-aligned like the multigrid levels),
  channel-major layout ``a[c, z, x, y]``, weights ``w[cout, cin, 27]`` with
  tap t = (dz+1)*9 + (dx+1)*3 + (dy+1), bias per output channel, optional
  ReLU;
* ``upsample``: trilinear node interpolation m+1 -> 2m+1 (the multigrid
  prolongation), and its transpose;
* ``mse``: mean squared error and its gradient;
* ``adam``: one Adam step (Kingma & Ba, 2015, Algorithm 1).

Pinning: ``conv3d`` against a brute-force loop on tiny inputs, its
backward against central finite differences, ``upsample_t`` against the
transpose of the assembled interpolation matrix (tests/test_nn_oracle.py).
"""
from __future__ import annotations

import numpy as np

OFFS = [(dz, dx, dy) for dz in (-1, 0, 1) for dx in (-1, 0, 1) for dy in (-1, 0, 1)]


def out_shape(shape, stride):
    if stride == 1:
        return tuple(shape)
    return tuple((n - 1) // 2 + 1 for n in shape)


def _pad(a):
    return np.pad(a, ((0, 0), (1, 1), (1, 1), (1, 1)))


def conv3d(a, w, b, stride=1, relu=False):
    """y[co, o] = sum_ci sum_t w[co, ci, t] * a[ci, stride*o + off_t] + b[co]."""
    cin, nz, nx, ny = a.shape
    cout = w.shape[0]
    oz, ox, oy = out_shape((nz, nx, ny), stride)
    ap = _pad(a)
    y = np.zeros((cout, oz, ox, oy))
    for t, (dz, dx, dy) in enumerate(OFFS):
        sl = ap[:, 1 + dz: 1 + dz + stride * (oz - 1) + 1: stride,
                1 + dx: 1 + dx + stride * (ox - 1) + 1: stride,
                1 + dy: 1 + dy + stride * (oy - 1) + 1: stride]
        y += np.einsum("oc,czxy->ozxy", w[:, :, t], sl)
    y += b[:, None, None, None]
    return np.maximum(y, 0.0) if relu else y


def conv3d_backward(a, w, gy, stride=1, y=None):
    """Gradients (ga, gw, gb) of sum(gy * conv3d(a, w, b)); with ``y`` given
    (the ReLU output), gy is first masked by y > 0."""
    if y is not None:
        gy = gy * (y > 0)
    cin, nz, nx, ny = a.shape
    _, oz, ox, oy = gy.shape
    ap = _pad(a)
    gap = np.zeros_like(ap)
    gw = np.zeros_like(w)
    for t, (dz, dx, dy) in enumerate(OFFS):
        idx = (slice(None), slice(1 + dz, 1 + dz + stride * (oz - 1) + 1, stride),
               slice(1 + dx, 1 + dx + stride * (ox - 1) + 1, stride),
               slice(1 + dy, 1 + dy + stride * (oy - 1) + 1, stride))
        gw[:, :, t] = np.einsum("ozxy,czxy->oc", gy, ap[idx])
        gap[idx] += np.einsum("oc,ozxy->czxy", w[:, :, t], gy)
    gb = gy.sum(axis=(1, 2, 3))
    return gap[:, 1:-1, 1:-1, 1:-1], gw, gb


def _interp_w(nc):
    nf = 2 * (nc - 1) + 1
    P = np.zeros((nf, nc))
    for i in range(nf):
        if i % 2 == 0:
            P[i, i // 2] = 1.0
        else:
            P[i, i // 2] = P[i, i // 2 + 1] = 0.5
    return P


def upsample(a):
    """Trilinear node interpolation, channel-wise: (C, m+1, ...) -> (C, 2m+1, ...)."""
    c, nz, nx, ny = a.shape
    Pz, Px, Py = _interp_w(nz), _interp_w(nx), _interp_w(ny)
    return np.einsum("Zz,Xx,Yy,czxy->cZXY", Pz, Px, Py, a)


def upsample_t(g):
    """Transpose of ``upsample`` (restriction with weights 1, 1/2, 1/4, 1/8)."""
    c, nz, nx, ny = g.shape
    Pz, Px, Py = _interp_w((nz - 1) // 2 + 1), _interp_w((nx - 1) // 2 + 1), _interp_w((ny - 1) // 2 + 1)
    return np.einsum("Zz,Xx,Yy,cZXY->czxy", Pz, Px, Py, g)


def mse(p, t):
    d = p - t
    return float(np.mean(d * d)), 2.0 * d / d.size


def adam(w, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v


# ---------------------------------------------------------------- network
# Layer table of the learned coarse correction (name, cin, cout, stride, relu);
# the same order as topoptlab_amd/learned.py, parameters packed per layer as
# [w (cout*cin*27), b (cout)].
LAYERS = [("E1", 6, 16, 1, True), ("E2", 16, 32, 2, True), ("E3", 32, 64, 2, True),
          ("E3b", 64, 64, 1, True), ("U2", 64, 32, 1, True), ("D2", 64, 32, 1, True),
          ("U1", 32, 16, 1, True), ("D1", 32, 16, 1, True), ("OUT", 16, 3, 1, False)]


def unpack(params):
    out, o = {}, 0
    for name, ci, co, _, _ in LAYERS:
        nw = co * ci * 27
        out[name] = (params[o:o + nw].reshape(co, ci, 27), params[o + nw:o + nw + co])
        o += nw + co
    return out


def net_forward(params, X0):
    P = unpack(params)
    cv = lambda n, a, s, r: conv3d(a, *P[n], stride=s, relu=r)
    A = {"X0": X0}
    A["S1"] = cv("E1", X0, 1, True)
    A["S2"] = cv("E2", A["S1"], 2, True)
    A["T3a"] = cv("E3", A["S2"], 2, True)
    A["T3b"] = cv("E3b", A["T3a"], 1, True)
    A["UP2"] = upsample(A["T3b"])
    A["V2"] = cv("U2", A["UP2"], 1, True)
    A["C2"] = np.concatenate([A["V2"], A["S2"]])
    A["T2"] = cv("D2", A["C2"], 1, True)
    A["UP1"] = upsample(A["T2"])
    A["V1"] = cv("U1", A["UP1"], 1, True)
    A["C1"] = np.concatenate([A["V1"], A["S1"]])
    A["T1"] = cv("D1", A["C1"], 1, True)
    A["Y"] = cv("OUT", A["T1"], 1, False)
    return A["Y"], A


def net_backward(params, A, gY):
    """Parameter gradient (packed like params) of sum(gY * Y)."""
    P = unpack(params)
    G = {}

    def bw(name, x, gy, s, y):
        ga, gw, gb = conv3d_backward(x, P[name][0], gy, s, y)
        G[name] = (gw, gb)
        return ga

    gT1 = bw("OUT", A["T1"], gY, 1, None)
    gC1 = bw("D1", A["C1"], gT1, 1, A["T1"])
    gUP1 = bw("U1", A["UP1"], gC1[:16], 1, A["V1"])
    gT2 = upsample_t(gUP1)
    gC2 = bw("D2", A["C2"], gT2, 1, A["T2"])
    gUP2 = bw("U2", A["UP2"], gC2[:32], 1, A["V2"])
    gT3b = upsample_t(gUP2)
    gT3a = bw("E3b", A["T3a"], gT3b, 1, A["T3b"])
    gS2 = gC2[32:] + bw("E3", A["S2"], gT3a, 2, A["T3a"])
    gS1 = gC1[16:] + bw("E2", A["S1"], gS2, 2, A["S2"])
    bw("E1", A["X0"], gS1, 1, A["S1"])
    return np.concatenate([np.concatenate([G[n][0].ravel(), G[n][1]]) for n, *_ in LAYERS])
