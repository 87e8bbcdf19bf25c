"""Optimality-criteria and MMA updates on the GPU.

``oc_top88`` keeps the signature of
topoptlab/optimizer/optimality_criterion.py:9-79 (Nguyen/Paulino top88
variant: bisection on lambda until (l2-l1)/(l1+l2) <= 1e-3, returning the new
design and gt).  The bisection runs on the device; each step is one fused
kernel (clamped update + compensated sum of dv*(xnew-x)).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, lib


def oc_top88(x, volfrac, dc, dv, g, el_flags=None, move=0.2, l1=0.0, l2=1e9):
    """``el_flags`` (nel integers or None): 1 marks a passive element (held
    at 0), 2 an active one (held at 1), set inside every bisection step as at
    optimality_criterion.py:68-71."""
    x = torch.as_tensor(x, dtype=torch.float64).reshape(-1).contiguous()
    dev = x.device
    dc = torch.as_tensor(dc, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    dv = torch.as_tensor(dv, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    if not (x.numel() == dc.numel() == dv.numel()):
        raise ValueError("x, dc and dv must have the same length")
    flags = None
    if el_flags is not None:
        flags = torch.as_tensor(el_flags, device=dev).reshape(-1).to(torch.int32).contiguous()
        if flags.numel() != x.numel():
            raise ValueError("el_flags must have one entry per element")
    xnew = torch.empty_like(x)
    gt = C.c_double()
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib().tl_oc_update(x.numel(), C.c_void_p(x.data_ptr()), C.c_void_p(dc.data_ptr()),
                             C.c_void_p(dv.data_ptr()),
                             C.c_void_p(0 if flags is None else flags.data_ptr()), float(g),
                             float(move), float(l1), float(l2), C.c_void_p(xnew.data_ptr()),
                             C.byref(gt), C.c_void_p(stream)), "tl_oc_update")
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


def mma_defaultkws(n: int, n_constr: int, ft=None, device=None):
    """Default MMA arguments of topoptlab/optimizer/mma_utils.py:8-60 as
    device tensors (xmin, xmax, low, upp: (n,)) and scalars."""
    if n_constr == 0:
        raise ValueError("GCMMA needs to have at least one constraint.")
    if n_constr != 1:
        raise NotImplementedError("the device MMA handles one constraint (the volume)")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else device
    kw = {"nconstr": n_constr,
          "xmin": torch.zeros(n, dtype=torch.float64, device=dev),
          "xmax": torch.ones(n, dtype=torch.float64, device=dev),
          "low": torch.ones(n, dtype=torch.float64, device=dev),
          "upp": torch.ones(n, dtype=torch.float64, device=dev),
          "a0": 1.0, "a": 0.0, "c": 10000.0, "d": 0.0,
          "move": 0.1 if ft in (5, 6) else (0.05 if ft == 7 else 0.2),
          "asyinit": 0.5, "asydecr": 0.7, "asyincr": 1.2, "asymin": 0.01, "asymax": 10.0,
          "raa0": 1e-5, "albefa": 0.1}
    return kw


def mma_update(x, xold1, xold2, df0dx, fval: float, dfdx, kw, iter: int):
    """One MMA step (the mmapy.mmasub call of topology_optimization.py:591-607
    for one constraint): returns the new design; kw["low"], kw["upp"] are
    updated in place, kw["newton_its"] holds the subproblem's Newton steps."""
    x = torch.as_tensor(x, dtype=torch.float64).reshape(-1).clone()
    dev = x.device
    vec = lambda v: None if v is None else \
        torch.as_tensor(v, dtype=torch.float64, device=dev).reshape(-1).contiguous()
    xo1, xo2, g0, g1 = vec(xold1), vec(xold2), vec(df0dx), vec(dfdx)
    opts = _lib.MmaOpts(*(float(kw[k]) for k in ("move", "asyinit", "asydecr", "asyincr",
                                                 "asymin", "asymax", "raa0", "albefa", "a0", "a",
                                                 "c", "d")))
    p = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
    its = C.c_int32()
    stream = torch.cuda.current_stream(dev).cuda_stream
    check(lib().tl_mma_update(x.numel(), p(x), p(xo1), p(xo2), p(kw.get("xmin")),
                              p(kw.get("xmax")), p(g0), float(fval), p(g1), p(kw["low"]),
                              p(kw["upp"]), int(iter), C.byref(opts), C.c_void_p(stream),
                              C.byref(its)), "tl_mma_update")
    kw["newton_its"] = its.value
    return x
