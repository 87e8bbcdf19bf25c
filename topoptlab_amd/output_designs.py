"""Design and field export (topoptlab/output_designs.py:36-179).

The reference writes through meshio, which is not a dependency here; this
module writes the same mesh and data as a legacy VTK unstructured grid
(binary big-endian by default, as meshio does, or ASCII):

* points: node (ix, iy[, iz]) at (ix, nely - iy[, iz]) times elem_size, in
  the mesh's node order (y fastest);
* cells: quad (VTK type 9) / hexahedron (12) with the element's nodes in the
  reference's local order (n1+1, n2+1, n2, n1[, n3+1, n4+1, n4, n3]);
* point data u{i}, u_bw{i}, f{i}, f_bw{i} (one per load case, ndpn
  components) and cell data xPhys, x, xTilde, xThresh, stress_vm, as FIELD
  arrays with the reference's names.

Device tensors are copied to the host first.
"""
from __future__ import annotations

from typing import Dict, Optional, Union

import numpy as np


def _host(a):
    if a is None:
        return None
    if hasattr(a, "detach"):
        a = a.detach().cpu().numpy()
    return np.asarray(a, dtype=np.float64)


def threshold(xPhys, volfrac: float) -> np.ndarray:
    """Black/white design with the volfrac*nel largest densities set to 1
    (output_designs.py:8-34)."""
    xPhys = _host(xPhys).reshape(-1, 1)
    indices = np.flip(np.argsort(xPhys[:, 0]))
    vt = int(np.floor(volfrac * xPhys.shape[0]))
    xThresh = np.zeros(xPhys.shape, order="F")
    xThresh[indices[:vt]] = 1.0
    print("Thresholded Vol.: {0:.5f}".format(vt / xThresh.shape[0]))
    return xThresh


def _mesh(nelx, nely, nelz, elem_size):
    if nelz is None:
        ix, iy = np.meshgrid(np.arange(nelx + 1), np.arange(nely + 1), indexing="ij")
        pts = np.column_stack([ix.ravel(), nely - iy.ravel(), np.zeros(ix.size)]).astype(float)
        ex, ey = np.meshgrid(np.arange(nelx), np.arange(nely), indexing="ij")
        n1 = ((nely + 1) * ex + ey).ravel()
        n2 = ((nely + 1) * (ex + 1) + ey).ravel()
        cells = np.column_stack([n1 + 1, n2 + 1, n2, n1])
        ctype, ndim = 9, 2
    else:
        iz, ix, iy = np.meshgrid(np.arange(nelz + 1), np.arange(nelx + 1), np.arange(nely + 1),
                                 indexing="ij")
        pts = np.column_stack([ix.ravel(), nely - iy.ravel(), iz.ravel()]).astype(float)
        ez, ex, ey = np.meshgrid(np.arange(nelz), np.arange(nelx), np.arange(nely), indexing="ij")
        pl = (nelx + 1) * (nely + 1)
        n1 = (pl * ez + (nely + 1) * ex + ey).ravel()
        n2 = (pl * ez + (nely + 1) * (ex + 1) + ey).ravel()
        n3 = (pl * (ez + 1) + (nely + 1) * ex + ey).ravel()
        n4 = (pl * (ez + 1) + (nely + 1) * (ex + 1) + ey).ravel()
        cells = np.column_stack([n1 + 1, n2 + 1, n2, n1, n3 + 1, n4 + 1, n4, n3])
        ctype, ndim = 12, 3
    if elem_size is not None:
        es = np.full(ndim, float(elem_size)) if np.isscalar(elem_size) else \
            np.asarray(elem_size, dtype=float)
        pts[:, :ndim] *= es[None, :]
    return pts, cells, ctype


def write_vtk(path: str, points: np.ndarray, cells: np.ndarray, cell_type: int,
              point_data: Dict[str, np.ndarray], cell_data: Dict[str, np.ndarray],
              binary: bool = True) -> None:
    """Legacy VTK 4.2 unstructured grid; data arrays as FIELD entries."""
    npts, ncell = points.shape[0], cells.shape[0]
    k = cells.shape[1]
    with open(path, "wb") as fh:
        def line(s):
            fh.write((s + "\n").encode())

        def block(a, dtype):
            a = np.ascontiguousarray(a)
            if binary:
                fh.write(a.astype(">" + dtype).tobytes())
                fh.write(b"\n")
            else:
                fmt = "%.17g" if dtype.startswith("f") else "%d"
                np.savetxt(fh, a.reshape(-1, a.shape[-1] if a.ndim > 1 else 1), fmt=fmt)
        line("# vtk DataFile Version 4.2")
        line("topoptlab_amd design export")
        line("BINARY" if binary else "ASCII")
        line("DATASET UNSTRUCTURED_GRID")
        line(f"POINTS {npts} double")
        block(points, "f8")
        line(f"CELLS {ncell} {ncell * (k + 1)}")
        block(np.column_stack([np.full(ncell, k), cells]).astype(np.int64), "i4")
        line(f"CELL_TYPES {ncell}")
        block(np.full((ncell, 1), cell_type), "i4")
        for kind, n, data in (("POINT_DATA", npts, point_data), ("CELL_DATA", ncell, cell_data)):
            if not data:
                continue
            line(f"{kind} {n}")
            line(f"FIELD FieldData {len(data)}")
            for name, arr in data.items():
                arr = np.asarray(arr, dtype=np.float64).reshape(n, -1)
                line(f"{name} {arr.shape[1]} {n} double")
                block(arr, "f8")


def export_vtk(filename: str, nelx: int, nely: int, xPhys, nelz: Optional[int],
               x=None, u=None, f=None, u_bw=None, f_bw=None, xTilde=None,
               elem_size: Union[None, float, np.ndarray] = None, volfrac: Optional[float] = None,
               stress_vm=None, binary: bool = True) -> None:
    """Write ``filename + ".vtk"`` (output_designs.py:36-179).  u, f, u_bw,
    f_bw: (ndof, nload) like the reference (or (nload, ndof) device tensors
    of this package); element arrays (nel,) or (nel, 1)."""
    points, cells, ctype = _mesh(nelx, nely, nelz, elem_size)
    npts, nel = points.shape[0], cells.shape[0]
    node_data = {}
    for name, arr in (("u", u), ("u_bw", u_bw), ("f", f), ("f_bw", f_bw)):
        arr = _host(arr)
        if arr is None:
            continue
        if arr.ndim == 1:
            arr = arr[:, None]
        if arr.shape[0] != npts * (arr.shape[0] // npts) or arr.shape[0] < npts:
            arr = arr.T  # (nload, ndof) -> (ndof, nload)
        for i in range(arr.shape[1]):
            node_data[f"{name}{i}"] = arr[:, i].reshape(npts, -1)
    el_data = {"xPhys": _host(xPhys).reshape(nel)}
    for name, arr in (("x", x), ("xTilde", xTilde)):
        if arr is not None:
            el_data[name] = _host(arr).reshape(nel)
    if volfrac is not None:
        el_data["xThresh"] = threshold(xPhys, volfrac)[:, 0]
    if stress_vm is not None:
        el_data["stress_vm"] = _host(stress_vm).reshape(nel, -1)
    write_vtk(filename + ".vtk", points, cells, ctype, node_data, el_data, binary=binary)
