"""Batched solves of many independent structured-grid problems on one GPU.

topoptlab solves one problem at a time and loops over the load cases of a
problem (``solve_linsystem.py:341-352``, ``topology_optimization.py:463-505``).
Here a whole batch of :class:`~topoptlab_amd.fem.StiffnessOperator` instances
of uneven grid sizes is handed to ``tl_batch_solve`` at once:

* every system (instance and load case) runs the same FP64 PCG as
  ``StiffnessOperator.solve`` (same kernels, same arithmetic, same stopping
  rule), so each result equals the single solve bit for bit;
* ``lockstep=True`` (2D instances with GMG V(1,1) cycles): all systems
  iterate together, each with its own PCG and V-cycle vectors; one PCG
  iteration of every system is recorded once, the recordings are merged, and
  each iteration is one launch per kernel and place in the cycle covering the
  blocks of every system (a tile table maps blocks to systems); converged
  systems drop out of the launches; with ``use_graphs`` each chunk of
  iterations is captured once per schedule (the set of systems still
  iterating) as a hipGraph on a stream of its own and replayed;
* otherwise instances run concurrently on several HIP streams; each
  instance's chunk of PCG iterations is captured once into a hipGraph and
  replayed until the device-side state reports convergence, and the load
  cases of one instance run one after another on its stream;
* ``batch_setup`` overlaps the instances' multigrid set-ups on several
  streams;
* ``autotune`` picks the 2D operator's workgroup shape per grid by timing.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, List, Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import check, lib
from .elements import lk_linear_elast_2d, lk_poisson_2d
from .example_bc import clamped_pointloads_2d, heatsink_2d
from .fem import PRECISIONS, PRECONDITIONERS, StiffnessOperator, auto_levels
from .workloads import gaussian_random_density

TILES = (0, 16, 32, 64)


def set_tile(op: StiffnessOperator, precision: str, ty: int) -> None:
    """Force the 2D operator's workgroup shape (0: node line; 16/32/64: tile height)."""
    check(lib().tl_set_tile2d(op._h, PRECISIONS[precision], int(ty)), "tl_set_tile2d")


def autotune(op: StiffnessOperator, reps: int = 20) -> Dict:
    """Time every 2D workgroup shape on this grid and keep the fastest."""
    tile = (C.c_int32 * 2)()
    us = (C.c_double * 8)()
    op._sync_stream()
    check(lib().tl_autotune(op._h, int(reps), tile, us), "tl_autotune")
    return {"fp64": tile[0], "fp32": tile[1],
            "us_fp64": dict(zip(TILES, us[0:4])), "us_fp32": dict(zip(TILES, us[4:8]))}


def clear_autotune_table() -> None:
    lib().tl_autotune_clear()


def batch_setup(ops: Sequence[StiffnessOperator], xs: Sequence, levels: Optional[Sequence[int]] = None,
                precision: str = "fp32", nstreams: int = 16, **mg_kw) -> None:
    """``set_density`` + ``setup_mg`` of every instance, issued on ``nstreams``
    side streams in turn so that the instances' set-up kernels overlap (each
    set-up is a chain of small dependent launches: the Galerkin coarsening and
    the ~60-90 Gauss-Jordan steps of its coarsest inverse).  The current
    stream waits for all of them before it returns; the results are those of
    the one-after-another set-up."""
    if len(xs) != len(ops) or (levels is not None and len(levels) != len(ops)):
        raise ValueError("need one density (and level count) per operator")
    main = torch.cuda.current_stream(ops[0].device) if ops else None
    streams = [torch.cuda.Stream(device=op.device) for op in ops[:max(1, nstreams)]]
    for s in streams:
        s.wait_stream(main)  # densities produced on the current stream
    for i, op in enumerate(ops):
        with torch.cuda.stream(streams[i % len(streams)]):
            op.set_density(xs[i])
            op.setup_mg(nlevels=None if levels is None else levels[i], precision=precision, **mg_kw)
    for s in streams:
        main.wait_stream(s)


def batch_solve(ops: Sequence[StiffnessOperator], fs: Sequence[torch.Tensor], rtol: float = 1e-8,
                atol: float = 0.0, maxiter: int = 100000, preconditioner: str = "gmg",
                nstreams: int = 16, chunk: int = 8, use_graphs: bool = True,
                x0: Optional[Sequence[torch.Tensor]] = None, raise_on_fail: bool = True,
                lockstep: bool = True):
    """Solve K_i u = f_i for every instance i; ``fs[i]`` holds one (ndof,) or
    several (nrhs, ndof) full-length float64 right-hand sides.  Returns the
    solutions (same shapes) and per instance a list of info dicts."""
    n = len(ops)
    if n == 0 or len(fs) != n:
        raise ValueError("need one right-hand-side tensor per operator")
    if preconditioner not in PRECONDITIONERS:
        raise ValueError(f"unknown preconditioner {preconditioner!r}")
    F, U, nrhs = [], [], []
    for i, (op, f) in enumerate(zip(ops, fs)):
        f = torch.as_tensor(f, dtype=torch.float64, device=op.device)
        squeeze = f.ndim == 1
        f = (f[None, :] if squeeze else f).contiguous()
        if f.shape[1] != op.ndof:
            raise ValueError(f"instance {i}: right-hand sides must have {op.ndof} entries")
        if preconditioner == "gmg" and op._mg_key is None:
            op.setup_mg()
        if x0 is not None:
            u = torch.as_tensor(x0[i], dtype=torch.float64, device=op.device).reshape(f.shape).clone()
        else:
            u = torch.empty_like(f)
        F.append(f)
        U.append((u, squeeze))
        nrhs.append(f.shape[0])
        op._sync_stream()
    handles = (C.c_void_p * n)(*[op._h.value for op in ops])
    fptr = (C.c_void_p * n)(*[f.data_ptr() for f in F])
    uptr = (C.c_void_p * n)(*[u.data_ptr() for u, _ in U])
    nr = (C.c_int32 * n)(*nrhs)
    infos = (_lib.SolveInfo * sum(nrhs))()
    opts = _lib.SolveOpts(float(rtol), float(atol), int(maxiter), int(x0 is not None),
                          PRECONDITIONERS[preconditioner], int(chunk))
    bo = _lib.BatchOpts(int(nstreams), int(chunk), int(bool(use_graphs)), int(bool(lockstep)))
    rc = lib().tl_batch_solve(n, handles, fptr, uptr, nr, C.byref(opts), C.byref(bo), infos)
    if rc != _lib.TL_ENOCONV or raise_on_fail:
        check(rc, "tl_batch_solve")
    out_info: List[List[Dict]] = []
    k = 0
    for i in range(n):
        lst = []
        for _ in range(nrhs[i]):
            inf = infos[k]
            lst.append({"iterations": inf.iterations, "status": inf.status,
                        "relres_recursive": inf.relres_recursive,
                        "relres_true": inf.relres_true})
            k += 1
        out_info.append(lst)
    sols = [u[0] if sq else u for u, sq in U]
    return sols, out_info


def build_instance(spec: Dict, device, max_coarse_dof: int = 3000):
    """StiffnessOperator, density and right-hand sides of one instance of
    ``workloads.batched_instances`` (E0=1, Emin=1e-9, nu=0.3 elasticity;
    k0=1, kmin=1e-3 heat; p=3)."""
    nelx, nely = spec["nelx"], spec["nely"]
    if spec["kind"] == "elast":
        KE, eps, ndpn = lk_linear_elast_2d(), 1e-9, 2
        ndof = 2 * (nelx + 1) * (nely + 1)
        _, f, fixed, _, _ = clamped_pointloads_2d(nelx, nely, ndof, load_nodes=spec["load_nodes"])
    else:
        KE, eps, ndpn = lk_poisson_2d(), 1e-3, 1
        ndof = (nelx + 1) * (nely + 1)
        _, f, fixed, _, _ = heatsink_2d(nelx, nely, ndof)
    x = gaussian_random_density(nelx, nely, seed=spec["density_seed"])
    op = StiffnessOperator(nelx, nely, KE=KE, fixed=fixed, eps=eps, device=device)
    levels = auto_levels(nelx, nely, None, ndpn, max_coarse_dof)
    return op, x, np.ascontiguousarray(f.T), levels
