"""Optimality-criteria update on the GPU.

``oc_top88`` keeps the signature of
topoptlab/optimizer/optimality_criterion.py:9-79 (Nguyen/Paulino top88
variant: bisection on lambda until (l2-l1)/(l1+l2) <= 1e-3, returning the new
design and gt).  The bisection runs on the device; each step is one fused
kernel (clamped update + compensated sum of dv*(xnew-x)).
"""
from __future__ import annotations

import ctypes as C

import torch

from ._lib import check, lib


def oc_top88(x, volfrac, dc, dv, g, el_flags=None, move=0.2, l1=0.0, l2=1e9):
    if el_flags is not None:
        raise NotImplementedError("element flags (passive/active elements) are not supported")
    x = torch.as_tensor(x, dtype=torch.float64).reshape(-1).contiguous()
    dev = x.device
    dc = torch.as_tensor(dc, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    dv = torch.as_tensor(dv, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    if not (x.numel() == dc.numel() == dv.numel()):
        raise ValueError("x, dc and dv must have the same length")
    xnew = torch.empty_like(x)
    gt = C.c_double()
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib().tl_oc_update(x.numel(), C.c_void_p(x.data_ptr()), C.c_void_p(dc.data_ptr()),
                             C.c_void_p(dv.data_ptr()), float(g), float(move), float(l1),
                             float(l2), C.c_void_p(xnew.data_ptr()), C.byref(gt),
                             C.c_void_p(stream)), "tl_oc_update")
    return xnew, gt.value


def mean_maxdiff(a, b=None):
    """(mean(a), max|a-b|) of two device vectors."""
    a = torch.as_tensor(a, dtype=torch.float64).reshape(-1).contiguous()
    bb = None if b is None else torch.as_tensor(b, dtype=torch.float64,
                                                device=a.device).reshape(-1).contiguous()
    m, d = C.c_double(), C.c_double()
    stream = torch.cuda.current_stream(a.device).cuda_stream
    check(lib().tl_mean_maxdiff(a.numel(), C.c_void_p(a.data_ptr()),
                                C.c_void_p(bb.data_ptr() if bb is not None else 0),
                                C.byref(m), C.byref(d), C.c_void_p(stream)), "tl_mean_maxdiff")
    return m.value, d.value
