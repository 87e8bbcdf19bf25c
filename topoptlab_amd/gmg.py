"""Host-side index helpers of the geometric multigrid hierarchy.

``create_coarse_inds`` / ``create_coarse_mask`` keep the signatures of
topoptlab/gmg.py:148-224: the fine DOF that coincides with each coarse DOF,
listed in the coarse grid's own numbering (DOF fastest, then y, x, z).  The
device hierarchy uses exactly this injection to derive coarse Dirichlet
masks (a coarse DOF is fixed iff its coincident fine DOF is fixed).
"""
from __future__ import annotations

from typing import Optional, Sequence, Union

import numpy as np


def _strides(stride, ndim):
    if isinstance(stride, (int, np.integer)):
        return (int(stride),) * ndim
    s = tuple(int(v) for v in stride)
    if len(s) != ndim:
        raise ValueError("stride must have one entry per dimension")
    return s


def create_coarse_inds(nelx: int, nely: int, nelz: Optional[int] = None, ndof: int = 1,
                       stride: Union[int, Sequence[int]] = 2) -> np.ndarray:
    ndim = 2 if nelz is None else 3
    sx, sy, *rest = _strides(stride, ndim)
    sz = rest[0] if rest else 1
    nx, ny = nelx + 1, nely + 1
    cx = np.arange(0, nx, sx)
    cy = np.arange(0, ny, sy)
    cz = np.arange(0, (nelz or 0) + 1, sz)
    # coarse numbering: y fastest, then x, then z; DOFs interleaved
    fz, fx, fy = np.meshgrid(cz, cx, cy, indexing="ij")
    node = (fz * nx + fx) * ny + fy
    return (ndof * node.ravel()[:, None] + np.arange(ndof)[None, :]).ravel()


def create_coarse_mask(nelx: int, nely: int, nelz: Optional[int] = None, ndof: int = 1,
                       stride: Union[int, Sequence[int]] = 2) -> np.ndarray:
    n = ndof * (nelx + 1) * (nely + 1) * (1 if nelz is None else nelz + 1)
    mask = np.zeros(n, dtype=bool)
    mask[create_coarse_inds(nelx, nely, nelz, ndof, stride)] = True
    return mask
