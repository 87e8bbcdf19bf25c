"""Distance (density/sensitivity) filter on the GPU.

``assemble_matrix_filter`` keeps the signature and return convention of
topoptlab/filter/matrix_filter.py:101-232 -- ``(H, Hs)`` used as
``H @ x / Hs`` or ``H @ (dc / Hs)`` -- but H is a ``DeviceFilter``: a stencil
operator that is never stored.  Its products accumulate each row in the
order of scipy's CSC product and without FMA contraction.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional

import torch

from ._lib import check, lib


class DeviceFilter:
    """H of the distance filter as a device operator (H @ v, H * v)."""

    def __init__(self, nelx: int, nely: int, rmin: float, nelz: Optional[int] = None,
                 device=None):
        self.ndim = 2 if nelz is None else 3
        self.nel = nelx * nely * (1 if nelz is None else nelz)
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            check(lib().tl_filter_create(self.ndim, nelx, nely, nelz or 0, float(rmin),
                                         C.c_void_p(self._stream()), C.byref(h)),
                  "tl_filter_create")
        self._h = h
        hs = torch.empty(self.nel, dtype=torch.float64, device=self.device)
        check(lib().tl_filter_hs(self._h, C.c_void_p(hs.data_ptr())), "tl_filter_hs")
        self.Hs = hs
        self.shape = (self.nel, self.nel)

    def _stream(self) -> int:
        return torch.cuda.current_stream(self.device).cuda_stream

    def close(self):
        if getattr(self, "_h", None):
            lib().tl_filter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def apply(self, v, mode: int = 3, x=None) -> torch.Tensor:
        """mode 0: H v / Hs ; 1: H (v / Hs) ; 2: H(x v)/Hs/max(1e-3, x) ; 3: H v."""
        shape = v.shape if isinstance(v, torch.Tensor) else None
        v = torch.as_tensor(v, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        if v.numel() != self.nel:
            raise ValueError(f"vector of length {v.numel()} for a filter of {self.nel} elements")
        out = torch.empty_like(v)
        xp = None
        if x is not None:
            xp = torch.as_tensor(x, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        check(lib().tl_filter_set_stream(self._h, C.c_void_p(self._stream())))
        check(lib().tl_filter_apply(self._h, C.c_void_p(v.data_ptr()), C.c_void_p(out.data_ptr()),
                                    int(mode), C.c_void_p(xp.data_ptr() if xp is not None else 0)),
              "tl_filter_apply")
        return out.reshape(shape) if shape is not None and len(shape) == 2 else out

    def __matmul__(self, v):
        return self.apply(v, 3)

    __mul__ = __matmul__


def assemble_matrix_filter(nelx: int, nely: int, rmin: float, nelz: Optional[int] = None,
                           pbc=False, device=None, **kwargs):
    """Return ``(H, Hs)`` with H a DeviceFilter and Hs its row sums (nel, 1)."""
    if pbc is not False and pbc is not None and bool(torch.as_tensor(pbc).any()):
        raise NotImplementedError("periodic filters are not supported")
    if not isinstance(rmin, (int, float)):
        raise NotImplementedError("only a scalar rmin is supported")
    H = DeviceFilter(nelx, nely, float(rmin), nelz, device)
    return H, H.Hs[:, None]


class HelmholtzFilter:
    """The Helmholtz PDE filter of topoptlab/filter/helmholtz_filter.py:16-130
    (Lazarov & Sigmund 2011, in the form of top88), on the GPU:
    x_filtered = TF^T KF^-1 TF x.

    KF is the screened-Poisson operator (k = (rmin / (2 sqrt 3))^2, no
    boundary conditions) applied matrix-free by a unit-scale scalar Q4
    ``StiffnessOperator``; KF^-1 is a PCG solve to ``rtol`` (the reference
    factorises KF with a sparse LU).  TF averages each element onto its 4
    nodes (``tl_helmholtz_transfer``).  2D only, like the reference (its 3D
    branch raises at helmholtz_filter.py:177)."""

    def __init__(self, nelx: int, nely: int, rmin: float, nelz: Optional[int] = None,
                 l=None, rtol: float = 1e-12, device=None, **kwargs):
        import numpy as np
        from .elements import lk_screened_poisson_2d
        from .fem import StiffnessOperator, auto_levels
        if nelz is not None:
            raise NotImplementedError("3D not yet completely implemented.")
        if l is not None and not np.allclose(np.asarray(l, dtype=float), 1.0):
            raise NotImplementedError("only unit elements are supported")
        self.nelx, self.nely = int(nelx), int(nely)
        self.nel = self.nelx * self.nely
        self.rtol = float(rtol)
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        Rmin = float(rmin) / (2.0 * np.sqrt(3.0))
        KE = lk_screened_poisson_2d(k=Rmin ** 2)
        # unit scale: eps = 0, penal = 1, density 1 gives s = 1 exactly
        self.KF = StiffnessOperator(self.nelx, self.nely, KE=KE, fixed=(), eps=0.0, penal=1.0,
                                    device=self.device)
        self.KF.set_density(torch.ones(self.nel, dtype=torch.float64, device=self.device))
        levels = auto_levels(self.nelx, self.nely, None, 1, 2000)
        self._prec = "gmg" if levels >= 2 else "jacobi"
        if levels >= 2:
            self.KF.setup_mg(nlevels=levels, precision="fp64")

    def close(self):
        if getattr(self, "KF", None) is not None:
            self.KF.close()
            self.KF = None

    def _transfer(self, v, mode):
        n_in = self.nel if mode == 0 else self.KF.ndof
        n_out = self.KF.ndof if mode == 0 else self.nel
        v = torch.as_tensor(v, dtype=torch.float64, device=self.device).reshape(-1).contiguous()
        if v.numel() != n_in:
            raise ValueError(f"vector of length {v.numel()}, expected {n_in}")
        out = torch.empty(n_out, dtype=torch.float64, device=self.device)
        check(lib().tl_helmholtz_transfer(self.nelx, self.nely, C.c_void_p(v.data_ptr()),
                                          C.c_void_p(out.data_ptr()), int(mode),
                                          C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)),
              "tl_helmholtz_transfer")
        return out

    def _filter(self, v):
        shape = v.shape if isinstance(v, torch.Tensor) else None
        xn = self._transfer(v, 0)
        y, _ = self.KF.solve(xn, rtol=self.rtol, preconditioner=self._prec)
        out = self._transfer(y, 1)
        return out.reshape(shape) if shape is not None and len(shape) == 2 else out

    def apply_filter(self, x, **kwargs):
        """TF^T KF^-1 TF x (helmholtz_filter.py:67-92)."""
        return self._filter(x)

    def apply_filter_dx(self, x_filtered, dx_filtered, **kwargs):
        """The same operator on the sensitivities (helmholtz_filter.py:94-130;
        the filter is symmetric)."""
        return self._filter(dx_filtered)
