"""Oracle: the reference's OC optimisation loop with a direct solve.

Test infrastructure only (see ``oracle/__init__.py``).  Restates the body of
``topoptlab/topology_optimization.py:main`` (loop at 411-673) for the branch
``optimizer="oc"``, ``filter_mode`` "matrix" or "helmholtz" (2D; branches at
360-365, 554-555, 568-570, 646-647), ``ft in {-1, 0, 1}``,
``lin_solver_kw={"name": "scipy-direct"}``, compliance objective, SIMP.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse.linalg import factorized, spsolve

from . import fem, mma


def topopt_loop(nelx, nely, volfrac, rmin, ft=1, nelz=None, bcs=fem.mbb_2d,
                KE=None, eps=1e-9, penal=3.0, nouteriter=2000,
                stop_on_change=True, history=None, filter_mode="matrix", optimizer="oc",
                el_flags=None):
    """Run the OC loop; return (x, obj) like ``topology_optimization.py:688``.

    ``history`` (a list) receives one (obj, vol, change) tuple per iteration,
    matching the log line of ``topology_optimization.py:669-670``.
    ``stop_on_change=False`` disables the ``change < 0.01`` break (672-673)
    so that a fixed number of iterations runs.  ``el_flags`` (OC only) is
    passed to ``oc_top88`` as at 579-581; the initial design is not changed
    for OC (the flag block at 248-259 belongs to the MMA branch).
    """
    if el_flags is not None and optimizer != "oc":
        raise NotImplementedError("element flags are restated for the OC branch only")
    ndim = 2 if nelz is None else 3
    n = int(np.prod([nelx, nely, nelz][:ndim]))
    if KE is None:
        KE = fem.ke_elast_2d() if ndim == 2 else fem.ke_elast_3d()
    ndpn = KE.shape[-1] // 2 ** ndim
    ndof = ndpn * int(np.prod(np.array([nelx, nely, nelz][:ndim]) + 1))
    edof = (fem.edofmat_2d(nelx, nely, ndpn) if ndim == 2
            else fem.edofmat_3d(nelx, nely, nelz, ndpn))
    x = volfrac * np.ones(n)
    xPhys = x.copy()
    g = 0.0
    helm = filter_mode == "helmholtz"
    if helm:
        KF, TF = fem.assemble_helmholtz_filter(nelx, nely, rmin)
        lu_solve = factorized(KF)
        hfilt = lambda v: TF.T @ lu_solve(TF @ v)
    else:
        H, Hs = fem.assemble_matrix_filter(nelx, nely, rmin, nelz)
        Hs = Hs[:, 0]
    if ndim == 2:
        f, fixed, free = bcs(nelx, nely, ndof)
    else:
        f, fixed, free = bcs(nelx, nely, nelz, ndof)
    u = np.zeros(f.shape)
    xold = x.copy()
    if optimizer == "mma":  # topology_optimization.py:244-246, mma_utils.py:8-60
        kw = mma.mma_defaultkws(n, 1, ft)
        xhist = [x[:, None].copy() for _ in range(3)]
    obj = 0.0
    for loop in range(nouteriter):
        scale = fem.simp(xPhys, eps, penal)
        K = fem.apply_bc(fem.assemble_K(KE, edof, scale, ndof), free)
        sol = spsolve(K, f[free])
        u[free, :] = sol.reshape(len(free), -1)
        obj, dc = fem.compliance_sens(KE, edof, u, xPhys, eps, penal)
        # volume constraint and its gradient (topology_optimization.py:513-520)
        dv = np.full(n, 1.0 / n) if optimizer == "mma" else np.ones(n)
        constr = xPhys.mean() - volfrac
        if ft == 0 and helm:
            dc = hfilt(dc * xPhys) / np.maximum(0.001, x)
        elif ft == 1 and helm:
            dc = hfilt(dc)
            dv = hfilt(dv)
        elif ft == 0:
            dc = (H @ (x * dc)) / Hs / np.maximum(0.001, x)
        elif ft == 1:
            dc = H @ (dc / Hs)
            dv = H @ (dv / Hs)
        xold = x.copy()
        if optimizer == "mma":
            # iter=0: the call site's leftover load-case index (see oracle/mma.py)
            out = mma.mmasub(1, n, 0, x[:, None], kw["xmin"], kw["xmax"], xhist[-1], xhist[-2],
                             obj, dc[:, None], np.array([[constr]]), dv[None, :], kw["low"],
                             kw["upp"], kw["a0"], kw["a"], kw["c"], kw["d"], move=kw["move"],
                             asyinit=kw["asyinit"], asydecr=kw["asydecr"], asyincr=kw["asyincr"],
                             asymin=kw["asymin"], asymax=kw["asymax"], raa0=kw["raa0"],
                             albefa=kw["albefa"])
            kw["low"], kw["upp"] = out[9], out[10]
            x = out[0][:, 0].copy()
            xhist = (xhist + [x[:, None].copy()])[-4:]
        else:
            x, g = fem.oc_top88(x, volfrac, dc, dv, g, el_flags=el_flags)
        if ft == 1 and helm:
            xPhys = hfilt(x)
        elif ft == 1:
            xPhys = (H @ x) / Hs
        else:
            xPhys = x.copy()
        change = np.abs(x - xold).max()
        if history is not None:
            history.append((obj, xPhys.mean(), change))
        if stop_on_change and change < 0.01:
            break
    return x, obj
