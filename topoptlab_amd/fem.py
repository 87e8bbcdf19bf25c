"""Device stiffness operator of a structured Q4/H8 grid (host wrapper of the
C ABI).

``StiffnessOperator`` stands where topoptlab holds the assembled, BC-reduced
scipy matrix ``K = apply_bc(assemble_matrix(...))``
(topoptlab/topology_optimization.py:419-451, fem.py:141-237): it represents
K_ff for the current density without ever assembling it.  ``K @ v`` accepts
either a free-DOF vector (length nfree, the reference's K_ff product) or a
full-length DOF vector (fixed entries ignored, returned as zero).

All vectors are float64 torch tensors on the operator's device.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib

PRECISIONS = {"fp64": _lib.TL_PREC_FP64, "fp32": _lib.TL_PREC_FP32,
              "bf16": _lib.TL_PREC_BF16, "fp16": _lib.TL_PREC_FP16}
PRECONDITIONERS = {None: _lib.TL_PRECOND_NONE, "none": _lib.TL_PRECOND_NONE,
                   "jacobi": _lib.TL_PRECOND_JACOBI, "gmg": _lib.TL_PRECOND_GMG}


def _stream_ptr(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _dev_vec(v, device, n=None) -> torch.Tensor:
    t = torch.as_tensor(v, dtype=torch.float64, device=device)
    t = t.reshape(-1).contiguous()
    if n is not None and t.numel() != n:
        raise ValueError(f"vector of length {t.numel()} where {n} is expected")
    return t


def auto_levels(nelx, nely, nelz=None, ndpn=1, max_coarse_dof=4000, min_levels=2):
    """Deepest hierarchy whose grids halve exactly, stopping once the coarsest
    grid has at most ``max_coarse_dof`` DOFs."""
    dims = [nelx, nely] + ([] if nelz is None else [nelz])
    levels = 1
    while all(d % 2 == 0 and d >= 2 for d in dims):
        ndof = ndpn * int(np.prod(np.array(dims) + 1))
        if ndof <= max_coarse_dof and levels >= min_levels:
            break
        dims = [d // 2 for d in dims]
        levels += 1
    return max(levels, 1)


class StiffnessOperator:
    """K_ff of a structured grid, matrix-free on the GPU."""

    def __init__(self, nelx: int, nely: int, nelz: Optional[int] = None, KE=None,
                 fixed: Sequence[int] = (), eps: float = 1e-9, penal: float = 3.0,
                 device=None):
        self.ndim = 2 if nelz is None else 3
        self.nelx, self.nely, self.nelz = int(nelx), int(nely), (None if nelz is None else int(nelz))
        KE = np.ascontiguousarray(np.asarray(KE, dtype=np.float64))
        nn = 2 ** self.ndim
        if KE.ndim != 2 or KE.shape[0] != KE.shape[1] or KE.shape[0] % nn:
            raise ValueError("KE must be a square (2^ndim * ndpn) matrix")
        self.KE = KE
        self.ndpn = KE.shape[0] // nn
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        self.fixed = np.unique(np.asarray(fixed, dtype=np.int64))
        self.eps, self.penal = float(eps), float(penal)
        nnodes = (self.nelx + 1) * (self.nely + 1) * (1 if nelz is None else self.nelz + 1)
        self.ndof = self.ndpn * nnodes
        self.nel = self.nelx * self.nely * (1 if nelz is None else self.nelz)
        self.free = np.setdiff1d(np.arange(self.ndof), self.fixed)
        self._free_t = torch.as_tensor(self.free, device=self.device)
        fx = np.ascontiguousarray(self.fixed)
        desc = _lib.GridDesc(self.ndim, self.nelx, self.nely, self.nelz or 0, self.ndpn,
                             KE.ctypes.data_as(C.POINTER(C.c_double)), self.eps, self.penal,
                             fx.ctypes.data_as(C.POINTER(C.c_int64)), len(fx),
                             C.c_void_p(_stream_ptr(self.device)))
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().tl_create(C.byref(desc), C.byref(h)), "tl_create")
        self._h = h
        self._mg_key = None
        self.last_info = None

    # -- lifetime ---------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            lib().tl_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _sync_stream(self):
        check(lib().tl_set_stream(self._h, C.c_void_p(_stream_ptr(self.device))))

    @property
    def shape(self):
        n = len(self.free)
        return (n, n)

    # -- density ------------------------------------------------------------
    def set_density(self, xPhys) -> None:
        x = _dev_vec(xPhys, self.device, self.nel)
        self._x = x  # keep alive while kernels run
        self._sync_stream()
        check(lib().tl_set_density(self._h, C.c_void_p(x.data_ptr())), "tl_set_density")
        self._mg_key = None

    # -- products -------------------------------------------------------------
    def full(self, v_free: torch.Tensor) -> torch.Tensor:
        out = torch.zeros(self.ndof, dtype=torch.float64, device=self.device)
        out[self._free_t] = _dev_vec(v_free, self.device, len(self.free))
        return out

    def matvec_full(self, u, precision: str = "fp64") -> torch.Tensor:
        u = _dev_vec(u, self.device, self.ndof)
        y = torch.empty_like(u)
        self._sync_stream()
        check(lib().tl_apply_K(self._h, C.c_void_p(u.data_ptr()), C.c_void_p(y.data_ptr()),
                               PRECISIONS[precision]), "tl_apply_K")
        return y

    def __matmul__(self, v):
        v = torch.as_tensor(v, dtype=torch.float64, device=self.device)
        if v.ndim == 2:
            return torch.stack([self @ v[:, i] for i in range(v.shape[1])], dim=1)
        if v.numel() == self.ndof:
            return self.matvec_full(v)
        if v.numel() == len(self.free):
            return self.matvec_full(self.full(v))[self._free_t]
        raise ValueError(f"vector of length {v.numel()} does not match K ({self.ndof} DOFs, "
                         f"{len(self.free)} free)")

    def diagonal(self) -> torch.Tensor:
        d = torch.empty(self.ndof, dtype=torch.float64, device=self.device)
        self._sync_stream()
        check(lib().tl_diag_K(self._h, C.c_void_p(d.data_ptr())), "tl_diag_K")
        return d[self._free_t]

    # -- multigrid ------------------------------------------------------------
    def setup_mg(self, nlevels: Optional[int] = None, omega: float = 0.6, nu1: int = 1,
                 nu2: int = 1, precision: str = "fp32", max_coarse_dof: int = 4000,
                 gamma: int = 1, gamma_from: int = 1) -> int:
        if nlevels is None:
            nlevels = auto_levels(self.nelx, self.nely, self.nelz, self.ndpn, max_coarse_dof)
        opts = _lib.MgOpts(int(nlevels), float(omega), int(nu1), int(nu2), PRECISIONS[precision],
                           int(gamma), int(gamma_from))
        self._sync_stream()
        check(lib().tl_setup_mg(self._h, C.byref(opts)), "tl_setup_mg")
        self._mg_key = (nlevels, omega, nu1, nu2, precision, gamma, gamma_from)
        return nlevels

    # -- solve ------------------------------------------------------------------
    def solve(self, f, u0=None, rtol: float = 1e-5, atol: float = 0.0, maxiter: int = 100000,
              preconditioner: Optional[str] = "gmg", check_every: int = 8, raise_on_fail=True):
        """PCG on K_ff u = f with full-length DOF vectors; returns (u, info)."""
        if preconditioner not in PRECONDITIONERS:
            raise ValueError(f"unknown preconditioner {preconditioner!r}")
        f = _dev_vec(f, self.device, self.ndof)
        if u0 is None:
            u = torch.zeros_like(f)
            given = 0
        else:
            u = _dev_vec(u0, self.device, self.ndof).clone()
            given = 1
        if preconditioner == "gmg" and self._mg_key is None:
            self.setup_mg()
        opts = _lib.SolveOpts(float(rtol), float(atol), int(maxiter), given,
                              PRECONDITIONERS[preconditioner], int(check_every))
        info = _lib.SolveInfo()
        self._sync_stream()
        rc = lib().tl_solve(self._h, C.c_void_p(f.data_ptr()), C.c_void_p(u.data_ptr()),
                            C.byref(opts), C.byref(info))
        self.last_info = {"iterations": info.iterations, "status": info.status,
                          "relres_recursive": info.relres_recursive,
                          "relres_true": info.relres_true}
        if rc == _lib.TL_ENOCONV and not raise_on_fail:
            return u, self.last_info
        check(rc, "tl_solve")
        return u, self.last_info

    # -- objective ------------------------------------------------------------
    def compliance_sens(self, u, xPhys):
        """Compliance and dc/dx summed over the load cases in the columns (or
        rows, for a [ncases][ndof] tensor) of u."""
        u = torch.as_tensor(u, dtype=torch.float64, device=self.device)
        if u.ndim == 1:
            u = u[None, :]
        elif u.shape[0] == self.ndof and u.shape[-1] != self.ndof:
            u = u.T
        u = u.contiguous()
        if u.shape[1] != self.ndof:
            raise ValueError("u must hold full-length DOF vectors")
        x = _dev_vec(xPhys, self.device, self.nel)
        dc = torch.empty(self.nel, dtype=torch.float64, device=self.device)
        obj = C.c_double()
        self._sync_stream()
        check(lib().tl_compliance_sens(self._h, C.c_void_p(u.data_ptr()), u.shape[0],
                                       C.c_void_p(x.data_ptr()), C.byref(obj),
                                       C.c_void_p(dc.data_ptr())), "tl_compliance_sens")
        return obj.value, dc
