"""Oracle: PCG, damped Jacobi and a geometric-multigrid V-cycle (FP64, scipy).

Test infrastructure only (see ``oracle/__init__.py``).

* ``pcg`` restates ``topoptlab/linear_solvers.py:50-136`` (stop on the
  recursively updated residual, ``||r|| < max(atol, rtol*||b||)``).
* ``jacobi`` restates one sweep of ``linear_solvers.py:306-372`` without its
  absolute early exit.
* ``vcycle`` restates ``topoptlab/multigrid.py:89-154`` as the textbook
  symmetric V(nu1, nu2): zero initial guess on every level (the reference
  restricts the fine iterate as coarse initial guess, ``multigrid.py:144``),
  Galerkin operators built once per hierarchy (the reference rebuilds them per
  call, ``multigrid.py:139-142``) and a direct solve on the coarsest level.
* ``prolongator`` builds the bi/trilinear stride-2 interpolation of
  ``topoptlab/gmg.py:34-146`` on the mesh's own node numbering (y fastest);
  it does not reproduce that function's x-fastest column indexing, which only
  agrees with the mesh numbering for square coarse grids.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, csr_matrix, diags
from scipy.sparse.linalg import spsolve


def _interp_1d(nf: int) -> csr_matrix:
    """1D linear interpolation from nf//2+1 coarse nodes to nf+1 fine nodes."""
    nc = nf // 2
    rows, cols, vals = [], [], []
    for i in range(nf + 1):
        if i % 2 == 0:
            rows.append(i); cols.append(i // 2); vals.append(1.0)
        else:
            rows += [i, i]; cols += [(i - 1) // 2, (i + 1) // 2]; vals += [0.5, 0.5]
    return coo_matrix((vals, (rows, cols)), shape=(nf + 1, nc + 1)).tocsr()


def prolongator(nelx, nely, nelz=None, ndpn=1) -> csr_matrix:
    """Fine-DOF x coarse-DOF interpolation matrix (stride 2 in every axis).

    Node numbering follows ``elements/bilinear_quadrilateral.py:7-48`` /
    ``elements/trilinear_hexahedron.py:7-58`` on both grids, so a node index
    is ((iz*(nx) + ix)*ny + iy) and the node-level operator is the Kronecker
    product Pz (x) Px (x) Py; DOFs are interleaved (ndpn per node).
    """
    from scipy.sparse import kron, identity
    if any(n % 2 for n in (nelx, nely) + (() if nelz is None else (nelz,))):
        raise ValueError("grid not divisible by the stride 2")
    Pn = kron(_interp_1d(nelx), _interp_1d(nely))
    if nelz is not None:
        Pn = kron(_interp_1d(nelz), Pn)
    return kron(Pn, identity(ndpn)).tocsr()


def coarse_fixed(nelx, nely, nelz, ndpn, fixed_fine):
    """Coarse DOFs fixed: those whose coincident fine DOF is fixed."""
    ndim = 2 if nelz is None else 3
    nf = np.array([nelx, nely, nelz][:ndim]) + 1
    nc = np.array([nelx, nely, nelz][:ndim]) // 2 + 1
    fixed = np.asarray(fixed_fine)
    node, comp = np.divmod(fixed, ndpn)
    if ndim == 2:
        ix, iy = np.divmod(node, nf[1])
        keep = (ix % 2 == 0) & (iy % 2 == 0)
        cn = (ix // 2) * nc[1] + iy // 2
    else:
        iz, rest = np.divmod(node, nf[0] * nf[1])
        ix, iy = np.divmod(rest, nf[1])
        keep = (ix % 2 == 0) & (iy % 2 == 0) & (iz % 2 == 0)
        cn = ((iz // 2) * nc[0] + ix // 2) * nc[1] + iy // 2
    return np.sort(ndpn * cn[keep] + comp[keep])


def check_aligned(P, fixed_fine, fixed_coarse) -> bool:
    """True iff every fixed fine DOF interpolates only from fixed coarse DOFs
    (then masking P on the fine side is implied by masking it on the coarse
    side, and nested and composite Galerkin hierarchies coincide)."""
    ncd = P.shape[1]
    cfree = np.ones(ncd, dtype=bool)
    cfree[fixed_coarse] = False
    sub = P[np.asarray(fixed_fine), :][:, np.flatnonzero(cfree)]
    return sub.nnz == 0 or np.all(sub.data == 0)


class Hierarchy:
    """Galerkin hierarchy on free DOFs: A_{l+1} = P_l^T A_l P_l with
    P_l = P[free_l][:, free_{l+1}]."""

    def __init__(self, A_ff, nelx, nely, nelz, ndpn, fixed, nlevels,
                 omega=0.6, nu1=1, nu2=1):
        self.A = [A_ff.tocsr()]
        self.P = []
        self.omega, self.nu1, self.nu2 = omega, nu1, nu2
        dims = [nelx, nely] + ([] if nelz is None else [nelz])
        ndim = len(dims)
        nn = int(np.prod(np.array(dims) + 1))
        free = np.setdiff1d(np.arange(ndpn * nn), fixed)
        for _ in range(nlevels - 1):
            P = prolongator(*dims[:2], dims[2] if ndim == 3 else None, ndpn)
            cfix = coarse_fixed(*dims[:2], dims[2] if ndim == 3 else None,
                                ndpn, fixed)
            if not check_aligned(P, fixed, cfix):
                raise ValueError("Dirichlet set not aligned with the coarse grid")
            cfree = np.setdiff1d(np.arange(P.shape[1]), cfix)
            Pl = P[free, :][:, cfree].tocsr()
            self.P.append(Pl)
            self.A.append((Pl.T @ self.A[-1] @ Pl).tocsr())
            dims = [d // 2 for d in dims]
            fixed, free = cfix, cfree
        self.Dinv = [1.0 / A.diagonal() for A in self.A]

    def vcycle(self, b, lvl=0, gamma=1, gamma_from=1):
        """V(nu1, nu2)-cycle from a zero guess; with gamma > 1 every level
        >= gamma_from repeats its cycle gamma times (W-cycle for gamma = 2):
        smooth, correct from a fresh coarse cycle, smooth, from the current x."""
        A, Dinv, w = self.A[lvl], self.Dinv[lvl], self.omega
        if lvl == len(self.A) - 1:
            return spsolve(A.tocsc(), b)
        x = np.zeros_like(b)
        for _ in range(gamma if lvl >= gamma_from else 1):
            for _ in range(self.nu1):
                x = x + w * Dinv * (b - A @ x)
            r = b - A @ x
            xc = self.vcycle(self.P[lvl].T @ r, lvl + 1, gamma, gamma_from)
            x = x + self.P[lvl] @ xc
            for _ in range(self.nu2):
                x = x + w * Dinv * (b - A @ x)
        return x


def jacobi(A, b, x, omega):
    """One damped-Jacobi sweep, ``linear_solvers.py:369``."""
    return x + omega * (b - A @ x) / A.diagonal()


def pcg(A, b, M, x0=None, rtol=1e-5, atol=0.0, maxiter=100000):
    """``linear_solvers.py:50-136``; M is a callable r -> z.

    Returns (x, info, iterations) with info 0 on convergence.
    """
    x = np.zeros_like(b) if x0 is None else x0.copy()
    tol = max(atol, rtol * np.linalg.norm(b))
    r = b - A @ x
    p = np.zeros_like(b)
    rho_prev = None
    for i in range(maxiter):
        if np.linalg.norm(r) < tol:
            return x, 0, i
        z = M(r)
        rho = r @ z
        p = z.copy() if i == 0 else z + (rho / rho_prev) * p
        q = A @ p
        alpha = rho / (p @ q)
        x += alpha * p
        r -= alpha * q
        rho_prev = rho
    return x, maxiter - 1, maxiter
