"""Oracle: numpy/scipy restatement of topoptlab's FE building blocks.

Test infrastructure only (see ``oracle/__init__.py``).  Paths in citations are
relative to the topoptlab repository root.
"""
from __future__ import annotations

import numpy as np
from scipy.sparse import coo_matrix, csc_matrix

# ---------------------------------------------------------------------------
# element DOF tables
# ---------------------------------------------------------------------------


def edofmat_2d(nelx: int, nely: int, ndpn: int) -> np.ndarray:
    """Q4 element DOF table.

    Follows ``topoptlab/elements/bilinear_quadrilateral.py:7-48``: node id
    (nely+1)*ix + iy, local node order (n1+1, n2+1, n2, n1) with n1 the
    element's node at (ix, iy) and n2 at (ix+1, iy); DOF = ndpn*node + d.
    Element id = nely*ix + iy.
    """
    out = np.empty((nelx * nely, 4 * ndpn), dtype=np.int64)
    for ix in range(nelx):
        for iy in range(nely):
            e = nely * ix + iy
            n1 = (nely + 1) * ix + iy
            n2 = (nely + 1) * (ix + 1) + iy
            nodes = (n1 + 1, n2 + 1, n2, n1)
            for a, n in enumerate(nodes):
                for d in range(ndpn):
                    out[e, a * ndpn + d] = ndpn * n + d
    return out


def edofmat_3d(nelx: int, nely: int, nelz: int, ndpn: int) -> np.ndarray:
    """H8 element DOF table.

    Follows ``topoptlab/elements/trilinear_hexahedron.py:7-58``: node id
    (nelx+1)(nely+1)*iz + (nely+1)*ix + iy, element id (iz*nelx + ix)*nely +
    iy, local order (n1+1, n2+1, n2, n1, n3+1, n4+1, n4, n3) where n3/n4 are
    n1/n2 on layer iz+1.
    """
    nxy = (nelx + 1) * (nely + 1)
    ez, ex, ey = np.meshgrid(np.arange(nelz), np.arange(nelx), np.arange(nely),
                             indexing="ij")
    ez, ex, ey = ez.ravel(), ex.ravel(), ey.ravel()
    n1 = nxy * ez + (nely + 1) * ex + ey
    n2 = nxy * ez + (nely + 1) * (ex + 1) + ey
    n3 = n1 + nxy
    n4 = n2 + nxy
    nodes = np.stack([n1 + 1, n2 + 1, n2, n1, n3 + 1, n4 + 1, n4, n3], axis=1)
    edof = (ndpn * nodes[:, :, None] + np.arange(ndpn)[None, None, :])
    return edof.reshape(nodes.shape[0], 8 * ndpn).astype(np.int64)


# ---------------------------------------------------------------------------
# element matrices
# ---------------------------------------------------------------------------

_GP = np.array([-1.0, 1.0]) / np.sqrt(3.0)
_REF2 = np.array([[-1, -1], [1, -1], [1, 1], [-1, 1]], dtype=float)
_REF3 = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], dtype=float)


def isotropic_2d(E=1.0, nu=0.3):
    """Plane-stress tensor, ``topoptlab/stiffness_tensors.py:148-177``."""
    return E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0],
                                          [0, 0, (1 - nu) / 2]])


def isotropic_3d(E=1.0, nu=0.3):
    """3D tensor exactly as ``topoptlab/stiffness_tensors.py:179-204`` writes it.

    Note: the shear diagonal is E(1-nu)/(2(1+nu)(1-2nu)), not the shear
    modulus E/(2(1+nu)).  The reference's closed-form H8 stiffness
    (``elements/linear_elasticity_3d.py:65-667``) is built from this same
    tensor (its K[0,1] = E/(24(1-2nu)) at l=1), so parity keeps it.
    """
    f = E / ((1 + nu) * (1 - 2 * nu))
    c = np.zeros((6, 6))
    c[:3, :3] = nu
    np.fill_diagonal(c[:3, :3], 1 - nu)
    c[3, 3] = c[4, 4] = c[5, 5] = (1 - nu) / 2
    return f * c


def ke_elast_2d(E=1.0, nu=0.3, l=1.0):
    """Q4 plane-stress stiffness by 2x2 Gauss quadrature.

    Follows ``elements/linear_elasticity_2d.py:9-67`` (quadrature form) with
    the unit element centred at the origin (nodes at +-l/2 in reference order
    ``elements/bilinear_quadrilateral.py:209-212``).
    """
    C = isotropic_2d(E, nu)
    K = np.zeros((8, 8))
    for xi in _GP:
        for eta in _GP:
            dN = np.stack([_REF2[:, 0] * (1 + eta * _REF2[:, 1]) / 4,
                           _REF2[:, 1] * (1 + xi * _REF2[:, 0]) / 4], axis=1)
            dNdx = dN * (2.0 / l)
            B = np.zeros((3, 8))
            B[0, 0::2] = dNdx[:, 0]
            B[1, 1::2] = dNdx[:, 1]
            B[2, 0::2] = dNdx[:, 1]
            B[2, 1::2] = dNdx[:, 0]
            K += B.T @ C @ B * (l / 2) ** 2
    return K


def ke_elast_2d_top88(E=1.0, nu=0.3):
    """Textbook top88 Q4 stiffness (Andreassen et al. 2011; Python form in
    ``monolithic_codes/topopt88_cholmod.py:190-212``), same node order."""
    k = np.array([1 / 2 - nu / 6, 1 / 8 + nu / 8, -1 / 4 - nu / 12,
                  -1 / 8 + 3 * nu / 8, -1 / 4 + nu / 12, -1 / 8 - nu / 8,
                  nu / 6, 1 / 8 - 3 * nu / 8])
    idx = [[0, 1, 2, 3, 4, 5, 6, 7], [1, 0, 7, 6, 5, 4, 3, 2],
           [2, 7, 0, 5, 6, 3, 4, 1], [3, 6, 5, 0, 7, 2, 1, 4],
           [4, 5, 6, 7, 0, 1, 2, 3], [5, 4, 3, 2, 1, 0, 7, 6],
           [6, 3, 4, 1, 2, 7, 0, 5], [7, 2, 1, 4, 3, 6, 5, 0]]
    return E / (1 - nu ** 2) * k[np.array(idx)]


def ke_elast_3d(E=1.0, nu=0.3, l=1.0):
    """H8 stiffness by 2x2x2 Gauss quadrature.

    Follows ``elements/linear_elasticity_3d.py:10-63`` with the tensor of
    ``isotropic_3d`` above and node order of
    ``elements/trilinear_hexahedron.py:252-259``; Voigt order
    (xx, yy, zz, yz, xz, xy).
    """
    C = isotropic_3d(E, nu)
    K = np.zeros((24, 24))
    R = _REF3
    for xi in _GP:
        for eta in _GP:
            for zeta in _GP:
                dN = np.stack([
                    R[:, 0] * (1 + eta * R[:, 1]) * (1 + zeta * R[:, 2]) / 8,
                    R[:, 1] * (1 + xi * R[:, 0]) * (1 + zeta * R[:, 2]) / 8,
                    R[:, 2] * (1 + xi * R[:, 0]) * (1 + eta * R[:, 1]) / 8],
                    axis=1) * (2.0 / l)
                B = np.zeros((6, 24))
                B[0, 0::3] = dN[:, 0]
                B[1, 1::3] = dN[:, 1]
                B[2, 2::3] = dN[:, 2]
                B[3, 1::3] = dN[:, 2]
                B[3, 2::3] = dN[:, 1]
                B[4, 0::3] = dN[:, 2]
                B[4, 2::3] = dN[:, 0]
                B[5, 0::3] = dN[:, 1]
                B[5, 1::3] = dN[:, 0]
                K += B.T @ C @ B * (l / 2) ** 3
    return K


def ke_poisson_2d(k=1.0):
    """Q4 Poisson stiffness, unit square: ``elements/poisson_2d.py:68-97`` at
    l=(1,1), g=0 evaluates to k*[[4,-1,-2,-1],...]/6."""
    return k * np.array([[4, -1, -2, -1], [-1, 4, -1, -2],
                         [-2, -1, 4, -1], [-1, -2, -1, 4]], dtype=float) / 6.0


def lm_mass_2d(p=1.0):
    """Q4 scalar mass matrix, unit square: the closed form of
    ``elements/mass_scalar_2d.py:68-95``."""
    return p * np.array([[1 / 9, 1 / 18, 1 / 36, 1 / 18], [1 / 18, 1 / 9, 1 / 18, 1 / 36],
                         [1 / 36, 1 / 18, 1 / 9, 1 / 18], [1 / 18, 1 / 36, 1 / 18, 1 / 9]])


def ke_screened_poisson_2d(k=1.0):
    """``elements/screenedpoisson_2d.py:9-36``: Poisson stiffness plus unit mass."""
    return ke_poisson_2d(k) + lm_mass_2d(1.0)


# ---------------------------------------------------------------------------
# material interpolation, assembly, boundary conditions
# ---------------------------------------------------------------------------


def simp(x, eps=1e-9, penal=3.0):
    """``topoptlab/material_interpolation.py:6-32``."""
    return eps + (1 - eps) * x ** penal


def simp_dx(x, eps=1e-9, penal=3.0):
    """``topoptlab/material_interpolation.py:34-62``."""
    return penal * (1 - eps) * x ** (penal - 1)


def assemble_K(KE, edof, scale, ndof):
    """COO assembly of sum_e scale_e*KE (``fem.py:141-178``, ``fem.py:239-271``
    "full" mode, ``topology_optimization.py:419-431``)."""
    nd = edof.shape[1]
    iK = np.tile(edof, nd).ravel()
    jK = np.repeat(edof, nd).ravel()
    sK = (KE[None, :, :] * np.asarray(scale).reshape(-1)[:, None, None]).ravel()
    return coo_matrix((sK, (iK, jK)), shape=(ndof, ndof)).tocsc()


def apply_bc(K, free):
    """Dirichlet-zero by deleting rows/cols, ``fem.py:229-231``."""
    return K[free, :][:, free]


def mbb_2d(nelx, nely, ndof):
    """``example_bc/lin_elast.py:6-46``."""
    f = np.zeros((ndof, 1))
    fixed = np.hstack((np.arange(0, 2 * (nely + 1), 2), np.array([ndof - 1])))
    f[1, 0] = -1
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


def cantilever_2d(nelx, nely, ndof):
    """``example_bc/lin_elast.py:100-139``."""
    f = np.zeros((ndof, 1))
    fixed = np.arange(0, 2 * (nely + 1))
    f[-1, 0] = -1
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


def cantilever_2d_twoloads(nelx, nely, ndof):
    """``example_bc/lin_elast.py:141-182``: two load cases."""
    f = np.zeros((ndof, 2))
    fixed = np.arange(0, 2 * (nely + 1))
    f[-1, 0] = -1
    f[2 * nelx * (nely + 1) + 1, 1] = 1
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


def cantilever_3d(nelx, nely, nelz, ndof, force_mode="linear"):
    """``example_bc/lin_elast.py:274-338``: wall at ix=0, z-symmetry on the
    iz=0 layer, downward load on the y-DOFs of nodes (ix=nelx, iy=nely) for
    all iz (the last node of each node plane)."""
    f = np.zeros((ndof, 1))
    nxy = (nelx + 1) * (nely + 1)
    fix = np.arange(0, 3 * (nely + 1), 3)
    fix = np.tile(fix, nelz + 1) + np.repeat(np.arange(0, ndof, 3 * nxy), nely + 1)
    fix = np.hstack((fix, fix + 1, fix + 2))
    zsym = np.arange(2, nxy * 3, 3)
    fixed = np.union1d(zsym, fix)
    if force_mode == "linear":
        f0 = -np.linspace(1.0, 0.0, nelz + 1)
    elif force_mode == "constant":
        f0 = -1.0
    else:
        raise ValueError(force_mode)
    f[np.arange(nxy * 3 - 2, ndof, nxy * 3), 0] = f0
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


def heatplate_2d(nelx, nely, ndof):
    """``example_bc/heat_conduction.py:5-49``."""
    f = np.full((ndof, 1), 0.01)
    fixed = np.arange(int(nely / 2 - nely / 20), int(nely / 2 + 1 + nely / 20))
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


def clamped_lineload_3d(nelx, nely, nelz, ndof):
    """The 3D cantilever of the 192x96x96 benchmark, stated from its text:
    every DOF of the ix=0 face fixed; a unit total downward (-y in the
    element frame, i.e. the y-DOF) line load on the bottom edge of the x=L
    face (nodes ix=nelx, iy=nely, iz=0..nelz), distributed consistently
    (1/nelz per interior node, 1/(2 nelz) at the two ends).  Not a reference
    function; the closest reference case is ``cantilever_3d``."""
    f = np.zeros((ndof, 1))
    nxy = (nelx + 1) * (nely + 1)
    iy, iz = np.meshgrid(np.arange(nely + 1), np.arange(nelz + 1), indexing="ij")
    nodes = (nxy * iz + iy).ravel()
    fixed = np.sort(np.concatenate([3 * nodes, 3 * nodes + 1, 3 * nodes + 2]))
    edge = nxy * np.arange(nelz + 1) + (nely + 1) * nelx + nely
    w = np.full(nelz + 1, 1.0 / nelz)
    w[0] = w[-1] = 0.5 / nelz
    f[3 * edge + 1, 0] = -w
    return f, fixed, np.setdiff1d(np.arange(ndof), fixed)


# ---------------------------------------------------------------------------
# density filter
# ---------------------------------------------------------------------------


def assemble_matrix_filter(nelx, nely, rmin, nelz=None):
    """Distance filter H and row sums Hs.

    Follows ``filter/matrix_filter.py:101-232`` (non-periodic branch): box
    half-width ceil(rmin)-1, weight max(0, ||rmin||_2/sqrt(ndim) - dist),
    explicit zeros kept, linear index nely*x + y + nelx*nely*z.
    """
    ndim = 2 if nelz is None else 3
    n = np.array([nelx, nely, nelz][:ndim])
    nel = int(np.prod(n))
    rvec = np.array(ndim * [float(rmin)])
    el = np.arange(nel)
    if ndim == 2:
        coords = np.column_stack(np.divmod(el, nely))
    else:
        z, rest = np.divmod(el, nelx * nely)
        x, y = np.divmod(rest, nely)
        coords = np.column_stack((x, y, z))
    dx = np.maximum(np.ceil(rvec) - 1, 0).astype(int)
    nb = np.column_stack([a.ravel() for a in
                          np.meshgrid(*[np.arange(-d, d + 1) for d in dx],
                                      indexing="ij")])
    r = np.linalg.norm(nb, axis=1, ord=2)
    nbc = coords[:, None, :] + nb[None, :, :]
    inside = np.all((nbc >= 0) & (nbc < n[None, None, :]), axis=2)
    lin = (nbc * np.array([nely, 1, nelx * nely][:ndim])[None, None, :]).sum(-1)
    iH = np.repeat(el, inside.sum(1))
    jH = lin[inside]
    sH = np.tile(np.linalg.norm(rvec, ord=2) / np.sqrt(ndim) - r, (nel, 1))[inside]
    sH = np.maximum(sH, 0.0)
    H = coo_matrix((sH, (iH, jH)), shape=(nel, nel)).tocsc()
    Hs = np.asarray(H.sum(1))
    return H, Hs


# ---------------------------------------------------------------------------
# objective, sensitivities, OC
# ---------------------------------------------------------------------------


def compliance_sens(KE, edof, u, xPhys, eps=1e-9, penal=3.0):
    """Compliance and its density sensitivity for all load cases.

    Follows ``objectives.py:7-55`` (ce = u_e.KE.u_e, obj += sum s_e ce,
    adjoint -u) and ``topology_optimization.py:463-505`` (dobj +=
    simp_dx * adj_e . (KE u_e) summed over cases).
    """
    x = np.asarray(xPhys).reshape(-1)
    obj = 0.0
    dc = np.zeros_like(x)
    for i in range(u.shape[1]):
        ue = u[edof, i]
        ke_u = ue @ KE.T
        ce = (ke_u * ue).sum(1)
        obj += (simp(x, eps, penal) * ce).sum()
        dc += simp_dx(x, eps, penal) * (-ue * ke_u).sum(1)
    return obj, dc


def oc_top88(x, volfrac, dc, dv, g, el_flags=None, move=0.2, l1=0.0, l2=1e9):
    """``optimizer/optimality_criterion.py:9-79``; ``el_flags`` 1 holds an
    element at 0 (passive), 2 at 1 (active), inside each step (68-71)."""
    xnew = np.zeros(x.shape)
    gt = g
    while (l2 - l1) / (l1 + l2) > 1e-3:
        lmid = 0.5 * (l2 + l1)
        xnew[:] = np.maximum(0.0, np.maximum(
            x - move, np.minimum(1.0, np.minimum(x + move,
                                                 x * np.sqrt(-dc / dv / lmid)))))
        if el_flags is not None:
            xnew[el_flags == 1] = 0
            xnew[el_flags == 2] = 1
        gt = g + np.sum(dv * (xnew - x))
        if gt > 0:
            l1 = lmid
        else:
            l2 = lmid
    return xnew, gt


def assemble_helmholtz_filter(nelx, nely, rmin):
    """``filter/helmholtz_filter.py:133-230`` (2D; the reference raises for
    3D at 177): KF sums the screened-Poisson element matrix with
    k = (rmin / (2 sqrt 3))^2 over all elements, no boundary conditions; TF
    (ndof x nel) maps each element's value to its 4 nodes with weight 1/4.
    The filter is x -> TF^T KF^-1 TF x (``helmholtz_filter.py:92``)."""
    from scipy.sparse import coo_matrix
    Rmin = rmin / (2.0 * np.sqrt(3.0))
    KE = ke_screened_poisson_2d(Rmin ** 2)
    edof = edofmat_2d(nelx, nely, 1)
    nel = nelx * nely
    ndof = (nelx + 1) * (nely + 1)
    KF = assemble_K(KE, edof, np.ones(nel), ndof).tocsc()
    TF = coo_matrix((np.full(4 * nel, 0.25), (edof.flatten(order="F"), np.tile(np.arange(nel), 4))),
                    shape=(ndof, nel)).tocsc()
    return KF, TF
