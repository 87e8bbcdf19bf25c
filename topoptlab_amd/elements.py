"""Element tables and element matrices of structured Q4/H8 grids (host side).

Signatures follow topoptlab:
* ``create_edofMat_2d``  <- topoptlab/elements/bilinear_quadrilateral.py:7-48
* ``create_edofMat_3d``  <- topoptlab/elements/trilinear_hexahedron.py:7-58
* ``lk_linear_elast_2d`` <- topoptlab/elements/linear_elasticity_2d.py:69-164
* ``lk_linear_elast_3d`` <- topoptlab/elements/linear_elasticity_3d.py:65-667
* ``lk_poisson_2d``      <- topoptlab/elements/poisson_2d.py:68-97

Element matrices are integrated with 2-point Gauss rules per axis, which is
exact for these bilinear/trilinear elements on rectangular (unsheared)
cells; only unit-aspect elements (l = 1, g = 0) are supported, the only case
the structured-grid kernels handle.
"""
from __future__ import annotations

import itertools

import numpy as np

# local node offsets (dx, dy, dz) from an element's (ix, iy, iz) corner node,
# in the reference node order (n1+1, n2+1, n2, n1, then layer iz+1)
LOCAL_OFFSETS_2D = np.array([(0, 1), (1, 1), (1, 0), (0, 0)])
LOCAL_OFFSETS_3D = np.array([(0, 1, 0), (1, 1, 0), (1, 0, 0), (0, 0, 0),
                             (0, 1, 1), (1, 1, 1), (1, 0, 1), (0, 0, 1)])


def create_edofMat_2d(nelx, nely, nnode_dof, dtype=np.int32, **kwargs):
    ix, iy = np.meshgrid(np.arange(nelx), np.arange(nely), indexing="ij")
    ix, iy = ix.ravel(), iy.ravel()
    n1 = (nely + 1) * ix + iy
    n2 = n1 + (nely + 1)
    nodes = (nely + 1) * (ix[:, None] + LOCAL_OFFSETS_2D[None, :, 0]) \
        + iy[:, None] + LOCAL_OFFSETS_2D[None, :, 1]
    edof = (nnode_dof * nodes[:, :, None] + np.arange(nnode_dof)[None, None, :])
    return (edof.reshape(len(ix), -1).astype(dtype), n1.astype(dtype), n2.astype(dtype),
            None, None)


def create_edofMat_3d(nelx, nely, nelz, nnode_dof, dtype=np.int32, **kwargs):
    iz, ix, iy = np.meshgrid(np.arange(nelz), np.arange(nelx), np.arange(nely), indexing="ij")
    iz, ix, iy = iz.ravel(), ix.ravel(), iy.ravel()
    nx, ny = nelx + 1, nely + 1
    o = LOCAL_OFFSETS_3D
    nodes = ((iz[:, None] + o[None, :, 2]) * nx + ix[:, None] + o[None, :, 0]) * ny \
        + iy[:, None] + o[None, :, 1]
    edof = (nnode_dof * nodes[:, :, None] + np.arange(nnode_dof)[None, None, :])
    n1 = (iz * nx + ix) * ny + iy
    n2 = n1 + ny
    n3 = n1 + nx * ny
    n4 = n2 + nx * ny
    return (edof.reshape(len(ix), -1).astype(dtype), n1.astype(dtype), n2.astype(dtype),
            n3.astype(dtype), n4.astype(dtype))


def _jacobian(l, g, ndim):
    """Edge vectors of the element as the columns of J (physical = J q for
    the unit cell coordinates q, element frame with y up): side lengths l and
    parallelogram angles g as the reference's closed forms use them
    (elements/linear_elasticity_2d.py:69-164, poisson_2d.py:68-97,
    poisson_3d.py): in 2D the y edge leans by l1 tan(g0) in x; in 3D the y
    edge by l1 tan(g0) and the z edge by l2 tan(g1), both in x."""
    lv = np.broadcast_to(np.asarray(l, dtype=float), (ndim,))
    gv = np.zeros(ndim - 1)
    gin = np.asarray(g, dtype=float).ravel()
    gv[:min(len(gin), ndim - 1)] = gin[:ndim - 1]
    J = np.diag(lv).astype(float)
    J[0, 1] = lv[1] * np.tan(gv[0])
    if ndim == 3:
        J[0, 2] = lv[2] * np.tan(gv[1])
    return J


def _gauss_grad(ndim, l=None, g=None):
    """Physical shape-function gradients at the 2^ndim Gauss points of the
    element with side lengths l and angles g (unit cube by default); returns
    list of (weight * det J, dN (nnode, ndim)).  The Jacobian is constant on
    a parallelogram/parallelepiped, so 2-point Gauss rules stay exact."""
    offs = LOCAL_OFFSETS_2D if ndim == 2 else LOCAL_OFFSETS_3D
    J = np.eye(ndim) if l is None else _jacobian(l, np.zeros(ndim - 1) if g is None else g, ndim)
    Jinv, detJ = np.linalg.inv(J), float(np.linalg.det(J))
    # reference-cell corner signs: the mesh index iy grows downwards while the
    # element's eta axis points up (elements/bilinear_quadrilateral.py:209-212)
    sgn = 2 * offs - 1
    sgn[:, 1] *= -1
    gp = 0.5 / np.sqrt(3.0)  # Gauss abscissa scaled to the unit cell
    out = []
    for q in itertools.product((-gp, gp), repeat=ndim):
        q = np.array(q)
        # N_i = prod_a (1/2 + s_ia x_a) on [-1/2,1/2]^ndim
        fac = 0.5 + sgn * q[None, :]
        dN = np.empty((len(offs), ndim))
        for a in range(ndim):
            others = np.prod(np.delete(fac, a, axis=1), axis=1)
            dN[:, a] = sgn[:, a] * others
        out.append((detJ / 2 ** ndim, dN @ Jinv))
    return out


def lk_linear_elast_2d(E=1.0, nu=0.3, plane_stress=True, l=np.array([1.0, 1.0]),
                       g=np.array([0.0]), t=1.0, **kwargs):
    if plane_stress:
        C = E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
    else:
        C = E / ((1 + nu) * (1 - 2 * nu)) * np.array([[1 - nu, nu, 0], [nu, 1 - nu, 0],
                                                       [0, 0, (1 - 2 * nu) / 2]])
    K = np.zeros((8, 8))
    for w, dN in _gauss_grad(2, l, g):
        B = np.zeros((3, 8))
        B[0, 0::2] = dN[:, 0]
        B[1, 1::2] = dN[:, 1]
        B[2, 0::2] = dN[:, 1]
        B[2, 1::2] = dN[:, 0]
        K += w * B.T @ C @ B
    return t * K


def lk_linear_elast_3d(E=1.0, nu=0.3, l=np.array([1.0, 1.0, 1.0]), g=np.array([0.0, 0.0]),
                       **kwargs):
    """H8 stiffness with the reference's isotropic tensor
    (topoptlab/stiffness_tensors.py:179-204), whose shear diagonal is
    E(1-nu)/(2(1+nu)(1-2nu)); the reference's closed form uses the same."""
    f = E / ((1 + nu) * (1 - 2 * nu))
    C = np.zeros((6, 6))
    C[:3, :3] = nu * f
    C[np.arange(3), np.arange(3)] = (1 - nu) * f
    C[np.arange(3, 6), np.arange(3, 6)] = (1 - nu) / 2 * f
    K = np.zeros((24, 24))
    for w, dN in _gauss_grad(3, l, g):
        B = np.zeros((6, 24))
        for a in range(3):
            B[a, a::3] = dN[:, a]
        B[3, 1::3], B[3, 2::3] = dN[:, 2], dN[:, 1]
        B[4, 0::3], B[4, 2::3] = dN[:, 2], dN[:, 0]
        B[5, 0::3], B[5, 1::3] = dN[:, 1], dN[:, 0]
        K += w * B.T @ C @ B
    return K


def lk_poisson_2d(k=1.0, l=np.array([1.0, 1.0]), g=np.array([0.0]), t=1.0, **kwargs):
    K = np.zeros((4, 4))
    for w, dN in _gauss_grad(2, l, g):
        K += w * dN @ dN.T
    return t * k * K


def lm_mass_2d(p=1.0, l=np.array([1.0, 1.0]), t=1.0, **kwargs):
    """Q4 scalar mass matrix p * int N N^T (elements/mass_scalar_2d.py:68-95),
    by 2x2 Gauss quadrature (exact for the biquadratic products), scaled by
    the element area l0*l1 as the reference does."""
    lv = np.broadcast_to(np.asarray(l, dtype=float), (2,))
    offs = LOCAL_OFFSETS_2D
    sgn = 2 * offs - 1
    sgn[:, 1] *= -1
    gp = 0.5 / np.sqrt(3.0)
    M = np.zeros((4, 4))
    for q in itertools.product((-gp, gp), repeat=2):
        N = np.prod(0.5 + sgn * np.array(q)[None, :], axis=1)
        M += 0.25 * np.outer(N, N)
    return p * t * lv[0] * lv[1] * M


def lk_screened_poisson_2d(k=1.0, l=np.array([1.0, 1.0]), g=np.array([0.0]), t=1.0, **kwargs):
    """Screened Poisson element matrix k * Poisson + mass
    (elements/screenedpoisson_2d.py:9-36); k is the squared filter radius in
    the Helmholtz filter."""
    return lk_poisson_2d(k=k, l=l, g=g, t=t) + lm_mass_2d(p=1.0, l=l, t=t)


def lk_poisson_3d(k=1.0, l=np.array([1.0, 1.0, 1.0]), g=np.array([0.0, 0.0]), **kwargs):
    K = np.zeros((8, 8))
    for w, dN in _gauss_grad(3, l, g):
        K += w * dN @ dN.T
    return k * K
