"""Linear-solver dispatch with the signature of topoptlab's ``solve_lin``
(topoptlab/solve_linsystem.py:20-365).

Solver names follow the reference's "library-solvername" scheme:

* ``"amd-pcg"``    -- device PCG (linear_solvers.py:50-136); preconditioner
  ``"amd-jacobi"`` (default) or ``None``.
* ``"amd-gmgpcg"`` -- device PCG preconditioned by the Galerkin GMG V-cycle
  (multigrid.py:89-223 with gmg.py's shape-function transfers).

``K`` is a :class:`topoptlab_amd.fem.StiffnessOperator` for the current
density; ``rhs``/``rhs0`` are (nfree, nbc) arrays or tensors.  Unknown solver
names raise ``ValueError("Unknown solver: ", solver)`` and non-convergence
raises ``RuntimeError`` as in the reference (349-351).

``preconditioner_kw`` of ``"amd-gmgpcg"`` takes the hierarchy options of
``StiffnessOperator.setup_mg`` (nlevels, omega, nu1, nu2, precision,
max_coarse_dof, gamma, gamma_from); other keys raise ``ValueError``.  The
returned ``P`` is the operator with its hierarchy; passing it back (the
adjoint solve of topology_optimization.py:479-487) reuses the hierarchy
while the density is unchanged.  ``solver_kw`` takes rtol, atol, maxiter and
check_every.
"""
from __future__ import annotations

from typing import Any, Dict, Optional

import torch

from .fem import StiffnessOperator
from .log_utils import EmptyLogger

SOLVERS = ("amd-pcg", "amd-gmgpcg")
MG_KEYS = ("nlevels", "omega", "nu1", "nu2", "precision", "max_coarse_dof", "gamma",
           "gamma_from")


def _solver_opts(solver_kw: Dict) -> Dict:
    kw = dict(solver_kw)
    out = {"rtol": kw.pop("rtol", 1e-5), "atol": kw.pop("atol", 0.0),
           "maxiter": kw.pop("maxiter", 100000), "check_every": kw.pop("check_every", 8)}
    if kw:
        raise ValueError("Unknown solver_kw: ", sorted(kw))
    return out


def solve_lin(K: StiffnessOperator, rhs, solver: str, rhs0=None, solver_kw: Dict = {},
              factorization=None, preconditioner: Optional[str] = None,
              preconditioner_kw: Dict = {}, P=None, logger=EmptyLogger, **kwargs: Any):
    if solver not in SOLVERS:
        raise ValueError("Unknown solver: ", solver)
    if not isinstance(K, StiffnessOperator):
        raise TypeError("K must be a topoptlab_amd StiffnessOperator")
    dev = K.device
    rhs = torch.as_tensor(rhs, dtype=torch.float64, device=dev)
    if rhs.ndim == 1:
        rhs = rhs[:, None]
    nfree = len(K.free)
    if rhs.shape[0] != nfree:
        raise ValueError(f"rhs has {rhs.shape[0]} rows, K has {nfree} free DOFs")
    x0 = None if rhs0 is None else torch.as_tensor(rhs0, dtype=torch.float64,
                                                   device=dev).reshape(nfree, -1)
    opts = _solver_opts(solver_kw)
    if solver == "amd-gmgpcg":
        prec = "gmg"
        mg_kw = {k: v for k, v in preconditioner_kw.items() if k != "name"}
        unknown = sorted(set(mg_kw) - set(MG_KEYS))
        if unknown:
            raise ValueError("Unknown preconditioner_kw: ", unknown)
        if P is None or K._mg_key is None:
            K.setup_mg(**mg_kw)
        P = K
    else:
        prec = {None: "none", "amd-jacobi": "jacobi", "jacobi": "jacobi",
                "none": "none"}.get(preconditioner, None)
        if prec is None:
            raise ValueError("Unknown preconditioner: ", preconditioner)
        if preconditioner is None and solver == "amd-pcg":
            prec = "jacobi"
    sol = torch.zeros((nfree, rhs.shape[1]), dtype=torch.float64, device=dev)
    infos = []
    for i in range(rhs.shape[1]):
        f = K.full(rhs[:, i])
        u0 = None if x0 is None else K.full(x0[:, i])
        u, info = K.solve(f, u0=u0, preconditioner=prec, raise_on_fail=False, **opts)
        infos.append(info)
        if info["status"] != 0:
            raise RuntimeError("pcg iteration did not converge for bc ", i)
        sol[:, i] = u[K._free_t]
        if not isinstance(logger, type):  # the default is the EmptyLogger class itself
            logger.perf(f"PCGit. {info['iterations']}")
    K.last_infos = infos
    return sol, None, P
