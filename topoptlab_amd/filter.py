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
