"""Learned coarse-grid correction for the FEM solve (no counterpart in
topoptlab).

A 3-level convolutional network (16/32/64 channels, fp16 weights, fp32
accumulation; every layer runs on the HIP kernels of ``tl_nn_*``) maps the
design and the load, sampled on the level-1 node grid of the multigrid
hierarchy, to the shape of the displacement there.  Inside PCG it provides
the coarse-grid part of the initial guess: the level-1 field is prolongated
to the fine grid and scaled by the energy-optimal factor
alpha = (f.x0)/(x0.K x0), so the guess is never worse than zero in the
energy norm.  The plain GMG-PCG stays the arbiter: the solve runs to the same
residual tolerance from the learned guess, and falls back to the zero guess
(the plain method) when the network is unavailable or the scaled guess is
degenerate.

Network (channel-major tensors [c, z, x, y], node grids G1 -> G2 -> G3 with
2m+1 -> m+1 nodes per level)::

    E1  conv 6->16          G1   relu  -> C1[16:32]  (skip)
    E2  conv 16->32 /2      G2   relu  -> C2[32:64]  (skip)
    E3  conv 32->64 /2      G3   relu  -> T3a
    E3b conv 64->64         G3   relu  -> T3b
    up(T3b) -> U2 conv 64->32  G2 relu  -> C2[0:32]
    D2  conv 64->32 (C2)    G2   relu  -> T2
    up(T2)  -> U1 conv 32->16  G1 relu  -> C1[0:16]
    D1  conv 32->16 (C1)    G1   relu  -> T1
    OUT conv 16->3          G1         -> Y
"""
from __future__ import annotations

import ctypes as C
import math
import time
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from ._lib import check, lib

NIN = 6  # input channels


def _p(t: Optional[torch.Tensor]):
    return C.c_void_p(0 if t is None else t.data_ptr())


def _down(shape):
    return tuple((n - 1) // 2 + 1 for n in shape)


@dataclass
class Layer:
    name: str
    cin: int
    cout: int
    stride: int
    relu: bool


LAYERS = [Layer("E1", NIN, 16, 1, True), Layer("E2", 16, 32, 2, True),
          Layer("E3", 32, 64, 2, True), Layer("E3b", 64, 64, 1, True),
          Layer("U2", 64, 32, 1, True), Layer("D2", 64, 32, 1, True),
          Layer("U1", 32, 16, 1, True), Layer("D1", 32, 16, 1, True),
          Layer("OUT", 16, 3, 1, False)]


class CoarseNet:
    """Parameters (fp32 master + fp16 copy + Adam moments in flat buffers)
    and the forward/backward passes built from the tl_nn_* kernels."""

    def __init__(self, device="cuda", seed: int = 0):
        self.device = torch.device(device)
        self.offsets: Dict[str, Tuple[int, int, int]] = {}
        n = 0
        for L in LAYERS:
            nw = L.cout * L.cin * 27
            self.offsets[L.name] = (n, nw, L.cout)
            n += nw + L.cout
        self.nparam = n
        g = np.random.default_rng(seed)
        w = np.zeros(n, np.float32)
        for L in LAYERS:
            o, nw, _ = self.offsets[L.name]
            std = math.sqrt(2.0 / (L.cin * 27)) * (1.0 if L.relu else 0.5)
            w[o:o + nw] = g.standard_normal(nw).astype(np.float32) * std
        self.w = torch.as_tensor(w, device=self.device)
        self.w16 = self.w.to(torch.float16)
        self.grad = torch.zeros_like(self.w)
        self.m = torch.zeros_like(self.w)
        self.v = torch.zeros_like(self.w)
        self.step = 0
        self._partials = torch.zeros(1024, dtype=torch.float64, device=self.device)
        self._loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._buf_shape = None

    # ---------------------------------------------------------------- params
    def _wb(self, name, half=True):
        o, nw, nb = self.offsets[name]
        src = self.w16 if half else self.w
        return src[o:o + nw], self.w[o + nw:o + nw + nb]

    def _gwb(self, name):
        o, nw, nb = self.offsets[name]
        return self.grad[o:o + nw], self.grad[o + nw:o + nw + nb]

    def state_dict(self):
        return {"w": self.w.detach().cpu().numpy().copy()}

    def load(self, w: np.ndarray):
        if w.size != self.nparam:
            raise ValueError(f"expected {self.nparam} parameters, got {w.size}")
        self.w.copy_(torch.as_tensor(w.astype(np.float32)))
        self.w16.copy_(self.w.to(torch.float16))

    # ---------------------------------------------------------------- buffers
    def _buffers(self, g1):
        if self._buf_shape == g1:
            return self.B
        g2, g3 = _down(g1), _down(_down(g1))
        d = self.device
        z = lambda c, g: torch.zeros((c,) + tuple(g), dtype=torch.float32, device=d)
        self.B = {"g1": g1, "g2": g2, "g3": g3,
                  "C1": z(32, g1), "C2": z(64, g2), "T3a": z(64, g3), "T3b": z(64, g3),
                  "UP2": z(64, g2), "T2": z(32, g2), "UP1": z(32, g1), "T1": z(16, g1),
                  "Y": z(3, g1),
                  "gC1": z(32, g1), "gC2": z(64, g2), "gT3a": z(64, g3), "gT3b": z(64, g3),
                  "gUP2": z(64, g2), "gT2": z(32, g2), "gUP1": z(32, g1), "gT1": z(16, g1),
                  "gY": z(3, g1)}
        self._buf_shape = g1
        return self.B

    # ---------------------------------------------------------------- kernels
    def _conv(self, name, x, g_in, y, stream):
        L = next(l for l in LAYERS if l.name == name)
        w16, b = self._wb(name)
        check(lib().tl_nn_conv3d(L.cin, L.cout, *g_in, L.stride, int(L.relu), _p(x), _p(w16),
                                 _p(b), _p(y), C.c_void_p(stream)), f"conv {name}")

    def _conv_bwd(self, name, x, g_in, gy, y_relu, gx, accumulate, stream):
        L = next(l for l in LAYERS if l.name == name)
        if L.relu:
            check(lib().tl_nn_relu_mask(gy.numel(), _p(y_relu), _p(gy), C.c_void_p(stream)), "relu")
        gw, gb = self._gwb(name)
        check(lib().tl_nn_conv3d_bwd_weight(L.cin, L.cout, *g_in, L.stride, _p(x), _p(gy), _p(gw),
                                            _p(gb), C.c_void_p(stream)), f"bwd weight {name}")
        if gx is not None:
            w16, _ = self._wb(name)
            check(lib().tl_nn_conv3d_bwd_data(L.cin, L.cout, *g_in, L.stride, _p(gy), _p(w16),
                                              _p(gx), int(accumulate), C.c_void_p(stream)),
                  f"bwd data {name}")

    def _up(self, c, g_coarse, x, y, stream, transpose=False, accumulate=False):
        fn = lib().tl_nn_upsample_t if transpose else lib().tl_nn_upsample
        check(fn(c, *g_coarse, _p(x), _p(y), int(accumulate), C.c_void_p(stream)), "upsample")

    # ---------------------------------------------------------------- passes
    def forward(self, X0: torch.Tensor) -> torch.Tensor:
        """X0: [6, nz1, nx1, ny1] fp32, node counts = 1 (mod 4) and >= 5 (two
        stride-2 levels on odd grids). Returns Y [3, ...]."""
        g1 = tuple(X0.shape[1:])
        if any(n % 4 != 1 or n < 5 for n in g1):
            raise ValueError(f"network grid needs node counts = 1 mod 4 and >= 5, got {g1}")
        B = self._buffers(g1)
        s = torch.cuda.current_stream(self.device).cuda_stream
        g2, g3 = B["g2"], B["g3"]
        self._X0 = X0
        self._conv("E1", X0, g1, B["C1"][16:], s)
        self._conv("E2", B["C1"][16:], g1, B["C2"][32:], s)
        self._conv("E3", B["C2"][32:], g2, B["T3a"], s)
        self._conv("E3b", B["T3a"], g3, B["T3b"], s)
        self._up(64, g3, B["T3b"], B["UP2"], s)
        self._conv("U2", B["UP2"], g2, B["C2"][:32], s)
        self._conv("D2", B["C2"], g2, B["T2"], s)
        self._up(32, g2, B["T2"], B["UP1"], s)
        self._conv("U1", B["UP1"], g1, B["C1"][:16], s)
        self._conv("D1", B["C1"], g1, B["T1"], s)
        self._conv("OUT", B["T1"], g1, B["Y"], s)
        return B["Y"]

    def backward(self, gY: torch.Tensor) -> None:
        """Parameter gradients into self.grad for dLoss/dY = gY (after forward)."""
        B = self.B
        s = torch.cuda.current_stream(self.device).cuda_stream
        g1, g2, g3 = B["g1"], B["g2"], B["g3"]
        B["gY"].copy_(gY)
        self._conv_bwd("OUT", B["T1"], g1, B["gY"], None, B["gT1"], False, s)
        self._conv_bwd("D1", B["C1"], g1, B["gT1"], B["T1"], B["gC1"], False, s)
        self._conv_bwd("U1", B["UP1"], g1, B["gC1"][:16], B["C1"][:16], B["gUP1"], False, s)
        self._up(32, g2, B["gUP1"], B["gT2"], s, transpose=True)
        self._conv_bwd("D2", B["C2"], g2, B["gT2"], B["T2"], B["gC2"], False, s)
        self._conv_bwd("U2", B["UP2"], g2, B["gC2"][:32], B["C2"][:32], B["gUP2"], False, s)
        self._up(64, g3, B["gUP2"], B["gT3b"], s, transpose=True)
        self._conv_bwd("E3b", B["T3a"], g3, B["gT3b"], B["T3b"], B["gT3a"], False, s)
        self._conv_bwd("E3", B["C2"][32:], g2, B["gT3a"], B["T3a"], B["gC2"][32:], True, s)
        self._conv_bwd("E2", B["C1"][16:], g1, B["gC2"][32:], B["C2"][32:], B["gC1"][16:], True, s)
        self._conv_bwd("E1", self._X0, g1, B["gC1"][16:], B["C1"][16:], None, False, s)

    def loss_and_grad(self, Y: torch.Tensor, T: torch.Tensor) -> torch.Tensor:
        """MSE(Y, T) as a device scalar; dLoss/dY into the gradient buffer."""
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().tl_nn_mse(Y.numel(), _p(Y), _p(T), _p(self.B["gY"]), _p(self._partials),
                              _p(self._loss), C.c_void_p(s)), "mse")
        return self._loss

    def adam_step(self, lr: float) -> None:
        self.step += 1
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().tl_nn_adam(self.nparam, _p(self.w), _p(self.grad), _p(self.m), _p(self.v),
                               _p(self.w16), self.step, C.c_float(lr), C.c_void_p(s)), "adam")

    def train_step(self, X0, T, lr=1e-3):
        Y = self.forward(X0)
        loss = self.loss_and_grad(Y, T).clone()
        self.backward(self.B["gY"].clone())
        self.adam_step(lr)
        return loss


# -------------------------------------------------------------------- features
def level1_shape(op) -> Tuple[int, int, int]:
    return (op.nelz // 2 + 1, op.nelx // 2 + 1, op.nely // 2 + 1)


def features(op, xPhys: torch.Tensor, f: torch.Tensor) -> torch.Tensor:
    """Network input [6, G1] for the operator's grid, design xPhys and load f."""
    fmax = float(f.abs().max().item()) if f.numel() else 0.0
    out = torch.empty((NIN,) + level1_shape(op), dtype=torch.float32, device=op.device)
    op._sync_stream()
    check(lib().tl_nn_features(op._h, _p(xPhys), _p(f), C.c_double(8.0 / fmax if fmax else 0.0),
                               _p(out)), "tl_nn_features")
    return out


def target(op, u: torch.Tensor) -> torch.Tensor:
    """Training target: the displacement at the level-1 nodes, max-normalised."""
    umax = float(u.abs().max().item())
    out = torch.empty((3,) + level1_shape(op), dtype=torch.float32, device=op.device)
    check(lib().tl_nn_inject(op._h, _p(u), C.c_double(1.0 / umax if umax else 0.0), _p(out)),
          "tl_nn_inject")
    return out


def learned_guess(net: CoarseNet, op, xPhys: torch.Tensor, f: torch.Tensor):
    """Prolongated, energy-scaled network prediction (ndof float64) and alpha."""
    Y = net.forward(features(op, xPhys, f))
    x0 = torch.empty(op.ndof, dtype=torch.float64, device=op.device)
    work = torch.empty_like(x0)
    check(lib().tl_nn_prolong(op._h, _p(Y), _p(x0)), "tl_nn_prolong")
    alpha = C.c_double()
    check(lib().tl_energy_scale(op._h, _p(f), _p(x0), _p(work), C.byref(alpha)), "tl_energy_scale")
    return x0, alpha.value


def solve_with_learned_guess(net: Optional[CoarseNet], op, xPhys, f, **solve_kw):
    """PCG from the learned initial guess; the plain zero-guess solve is the
    fallback (no network, or a degenerate/non-finite scaled guess).  Returns
    (u, info) with info["learned"] = True when the guess was used."""
    if net is not None:
        x0, alpha = learned_guess(net, op, xPhys, f)
        if math.isfinite(alpha) and alpha > 0.0:
            u, info = op.solve(f, u0=x0, **solve_kw)
            info = dict(info, learned=True, alpha=alpha)
            return u, info
    u, info = op.solve(f, **solve_kw)
    return u, dict(info, learned=False, alpha=0.0)


# -------------------------------------------------------------------- data
def cantilever_point_load(nelx, nely, nelz, rng):
    """x = 0 face clamped; unit downward point load at a seeded node of the
    x = nelx face (the seeded load positions of the training runs)."""
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    ndof = 3 * nx * ny * nz
    iz, iy = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    nodes = (iz * nx * ny + iy).ravel()
    fixed = np.sort((3 * nodes[:, None] + np.arange(3)[None, :]).ravel())
    jy, jz = int(rng.integers(0, ny)), int(rng.integers(0, nz))
    node = (jz * nx + nelx) * ny + jy
    f = np.zeros(ndof)
    f[3 * node + 1] = -1.0
    return f, fixed


def generate_pairs(nelx, nely, nelz, runs, iters, seed=0, volfrac=0.3, rmin=1.5,
                   device="cuda", log=None, designs=None):
    """(features, target) pairs from plain GMG-PCG solves along OC runs of
    cantilevers with seeded load positions (density filter, OC, move 0.2).
    With a list ``designs``, each run's last (xPhys, f, fixed) is appended."""
    return _oc_runs(nelx, nely, nelz, runs, iters, seed, volfrac, rmin, device, log, designs,
                    True)


def _oc_runs(nelx, nely, nelz, runs, iters, seed, volfrac, rmin, device, log, designs, pairs):
    from .elements import lk_linear_elast_3d
    from .fem import StiffnessOperator
    from .filter import DeviceFilter
    from .optimizer import oc_top88
    rng = np.random.default_rng(seed)
    KE = lk_linear_elast_3d()
    nel = nelx * nely * nelz
    flt = DeviceFilter(nelx, nely, rmin, nelz=nelz, device=device)
    X, T = [], []
    for run in range(runs):
        f_h, fixed = cantilever_point_load(nelx, nely, nelz, rng)
        op = StiffnessOperator(nelx, nely, nelz, KE=KE, fixed=fixed, device=device)
        f = torch.as_tensor(f_h, device=device)
        x = torch.full((nel,), volfrac, dtype=torch.float64, device=device)
        dv = torch.ones(nel, dtype=torch.float64, device=device)
        u = None
        for it in range(iters):
            xPhys = flt.apply(x, mode=0)
            op.set_density(xPhys)
            op.setup_mg(precision="fp32")
            u, info = op.solve(f, u0=u, rtol=1e-6)
            if pairs:
                X.append(features(op, xPhys, f))
                T.append(target(op, u))
            obj, dc = op.compliance_sens(u, xPhys)
            dc = flt.apply(dc, mode=1)
            dvf = flt.apply(dv, mode=1)
            x, _ = oc_top88(x, volfrac, dc, dvf, 0.0)
        if log:
            log(f"run {run}: {iters} OC iterations, last PCG {info['iterations']} its")
        if designs is not None:
            designs.append((xPhys.clone(), f, fixed))
        op.close()
    flt.close()
    return X, T


def oc_design(nelx, nely, nelz, seed=0, iters=8, volfrac=0.3, rmin=1.5, device="cuda"):
    """A design of the training distribution at any size: ``iters`` OC
    iterations (density filter, move 0.2) of a cantilever with a seeded point
    load, as ``generate_pairs`` runs them.  Returns (xPhys, f, fixed)."""
    designs = []
    _oc_runs(nelx, nely, nelz, 1, iters, seed, volfrac, rmin, device, None, designs, False)
    return designs[0]


def generate_field_pairs(nelx, nely, nelz, count, seed=0, device="cuda", log=None):
    """(features, target) pairs on random designs: seeded Gaussian random
    density fields (length scale 2-6 elements, mean 0.2-0.5, in [0.001, 1])
    under the clamped line load of the 3D cantilever benchmark or a seeded
    point load, solved by plain GMG-PCG (relres 1e-6)."""
    from .elements import lk_linear_elast_3d
    from .example_bc import clamped_lineload_3d
    from .fem import StiffnessOperator
    from .workloads import gaussian_random_density
    rng = np.random.default_rng(seed)
    KE = lk_linear_elast_3d()
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    X, T = [], []
    _, f_line, fixed_line, _, _ = clamped_lineload_3d(nelx, nely, nelz, ndof)
    for k in range(count):
        x_h = gaussian_random_density(nelx, nely, nelz, seed=int(rng.integers(1 << 30)),
                                      sigma=float(rng.uniform(2.0, 6.0)),
                                      mean=float(rng.uniform(0.2, 0.5)))
        if k % 2 == 0:
            f_h, fixed = f_line[:, 0], fixed_line
        else:
            f_h, fixed = cantilever_point_load(nelx, nely, nelz, rng)
        op = StiffnessOperator(nelx, nely, nelz, KE=KE, fixed=fixed, device=device)
        xPhys = torch.as_tensor(x_h, device=device)
        f = torch.as_tensor(np.ascontiguousarray(f_h), device=device)
        op.set_density(xPhys)
        op.setup_mg(precision="fp32")
        u, info = op.solve(f, rtol=1e-6)
        X.append(features(op, xPhys, f))
        T.append(target(op, u))
        op.close()
        if log and (k + 1) % 50 == 0:
            log(f"field pairs: {k + 1} of {count}")
    return X, T


def train(net: CoarseNet, X: List[torch.Tensor], T: List[torch.Tensor], epochs=20, lr=1e-3,
          seed=0, log=None) -> List[float]:
    """Adam over the pairs in seeded random order; returns the mean loss per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(X))
        losses = []
        for i in order:
            losses.append(net.train_step(X[i], T[i], lr))
        tot = float(torch.cat(losses).mean().item())
        hist.append(tot)
        if log:
            log(f"epoch {ep}: loss {tot:.4e}")
    return hist
