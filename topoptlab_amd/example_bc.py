"""Boundary-condition generators with the reference's signature
``bcs(nelx, nely, nelz, ndof) -> (u, f, fixed, free, springs)``.

* ``mbb_2d``        <- topoptlab/example_bc/lin_elast.py:6-46
* ``cantilever_2d`` <- topoptlab/example_bc/lin_elast.py:100-139
* ``cantilever_2d_twoloads`` <- topoptlab/example_bc/lin_elast.py:141-182
* ``cantilever_3d`` <- topoptlab/example_bc/lin_elast.py:274-338
* ``heatplate_2d``  <- topoptlab/example_bc/heat_conduction.py:5-49
* ``rectangle_2d``  <- topoptlab/example_bc/heat_conduction.py:106-145
* ``clamped_lineload_3d``: the 192x96x96 cantilever benchmark (not a
  reference function): x=0 face clamped, unit total downward line load on the
  bottom edge (iy = nely) of the x = L face, consistent nodal weights.
* ``clamped_pointloads_2d`` / ``heatsink_2d``: the batched-instance cases.
"""
from __future__ import annotations

import numpy as np


def _out(ndof, f, fixed):
    fixed = np.unique(np.asarray(fixed, dtype=np.int64))
    return (np.zeros(f.shape), f, fixed, np.setdiff1d(np.arange(ndof), fixed), None)


def mbb_2d(nelx, nely, ndof, **kwargs):
    f = np.zeros((ndof, 1))
    f[1, 0] = -1.0
    fixed = np.concatenate([2 * np.arange(nely + 1), [ndof - 1]])
    return _out(ndof, f, fixed)


def cantilever_2d(nelx, nely, ndof, **kwargs):
    f = np.zeros((ndof, 1))
    f[ndof - 1, 0] = -1.0
    return _out(ndof, f, np.arange(2 * (nely + 1)))


def cantilever_2d_twoloads(nelx, nely, ndof, **kwargs):
    """topoptlab/example_bc/lin_elast.py:141-182: two load cases."""
    f = np.zeros((ndof, 2))
    f[ndof - 1, 0] = -1.0
    f[2 * (nely + 1) * nelx + 1, 1] = 1.0
    return _out(ndof, f, np.arange(2 * (nely + 1)))


def cantilever_3d(nelx, nely, nelz, ndof, force_mode="linear", **kwargs):
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    iz, iy = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    wall = (iz * nx * ny + iy).ravel()
    fixed = np.concatenate([3 * wall, 3 * wall + 1, 3 * wall + 2,
                            3 * np.arange(nx * ny) + 2])
    f = np.zeros((ndof, 1))
    tip = np.arange(nz) * nx * ny + (nx - 1) * ny + (ny - 1)  # ix = nelx, iy = nely
    if force_mode == "linear":
        f[3 * tip + 1, 0] = -np.linspace(1.0, 0.0, nz)
    elif force_mode == "constant":
        f[3 * tip + 1, 0] = -1.0
    else:
        raise ValueError("force_mode must be either 'linear' or 'constant': ", force_mode)
    return _out(ndof, f, fixed)


def heatplate_2d(nelx, nely, ndof, **kwargs):
    f = np.full((ndof, 1), 0.01)
    fixed = np.arange(int(nely / 2 - nely / 20), int(nely / 2 + 1 + nely / 20))
    return _out(ndof, f, fixed)


def rectangle_2d(nelx, nely, ndof, **kwargs):
    """Heat conduction through a rectangular beam: the nodes of the x = 0
    edge held at zero (the heat sink), a unit source on every node of the
    x = L edge (the last nely+1 DOFs in the node-major numbering)."""
    f = np.zeros((ndof, 1))
    f[ndof - (nely + 1):, 0] = 1.0
    return _out(ndof, f, np.arange(nely + 1))


def clamped_lineload_3d(nelx, nely, nelz, ndof, **kwargs):
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    iz, iy = np.meshgrid(np.arange(nz), np.arange(ny), indexing="ij")
    wall = (iz * nx * ny + iy).ravel()
    fixed = np.concatenate([3 * wall, 3 * wall + 1, 3 * wall + 2])
    f = np.zeros((ndof, 1))
    edge = np.arange(nz) * nx * ny + (nx - 1) * ny + (ny - 1)
    w = np.full(nz, 1.0 / nelz)
    w[[0, -1]] = 0.5 / nelz
    f[3 * edge + 1, 0] = -w
    return _out(ndof, f, fixed)


def clamped_pointloads_2d(nelx, nely, ndof, load_nodes=(), **kwargs):
    """Left edge clamped; one load case per node in ``load_nodes`` with a unit
    downward point load."""
    fixed = np.arange(2 * (nely + 1))
    f = np.zeros((ndof, max(1, len(load_nodes))))
    for i, n in enumerate(load_nodes):
        f[2 * int(n) + 1, i] = -1.0
    return _out(ndof, f, fixed)


def heatsink_2d(nelx, nely, ndof, **kwargs):
    """Uniform unit heat source, T = 0 on the left edge (x = 0)."""
    f = np.full((ndof, 1), 1.0 / ((nelx + 1) * (nely + 1)))
    return _out(ndof, f, np.arange(nely + 1))
