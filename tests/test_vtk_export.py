"""CPU checks of the VTK export (output_designs.py:36-179): the file is
read back with an independent parser of the legacy VTK format and the mesh
is checked against element geometry (each quad/hexahedron spans a unit cell
with positive orientation in the reference's node order); data arrays come
back unchanged in both encodings.  No VTK reader library is installed here,
so byte-level parity with meshio's writer is unpinned."""
import numpy as np
import pytest

from topoptlab_amd.output_designs import export_vtk, threshold


def read_vtk(path):
    raw = open(path, "rb").read()
    pos = 0

    def getline():
        nonlocal pos
        e = raw.index(b"\n", pos)
        s = raw[pos:e].decode()
        pos = e + 1
        return s

    hdr = [getline() for _ in range(4)]
    binary = hdr[2] == "BINARY"

    def block(count, dtype):
        nonlocal pos
        if binary:
            size = np.dtype(dtype).itemsize * count
            a = np.frombuffer(raw[pos:pos + size], dtype=">" + dtype[1:] if dtype[0] == "<" else dtype)
            pos += size + 1
            return a.astype(np.float64 if "f" in dtype else np.int64)
        vals = []
        while len(vals) < count:
            vals += getline().split()
        return np.array(vals, dtype=float if "f" in dtype else np.int64)

    out = {"point_data": {}, "cell_data": {}}
    n, _ = getline().split()[1:3]
    out["points"] = block(int(n) * 3, "<f8").reshape(-1, 3)
    _, nc, size = getline().split()
    c = block(int(size), "<i4").reshape(int(nc), -1)
    out["cells"] = c[:, 1:]
    _, nt = getline().split()
    out["types"] = block(int(nt), "<i4")
    while pos < len(raw):
        l = getline().split()
        if not l:
            continue
        if l[0] in ("POINT_DATA", "CELL_DATA"):
            tgt = out["point_data" if l[0] == "POINT_DATA" else "cell_data"]
            nf = int(getline().split()[2])
            for _ in range(nf):
                name, nco, ntu, _ = getline().split()
                tgt[name] = block(int(nco) * int(ntu), "<f8").reshape(int(ntu), int(nco))
    return out


@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("dims", [(5, 3, None), (4, 3, 2)])
def test_export_vtk_roundtrip(tmp_path, binary, dims):
    nelx, nely, nelz = dims
    ndim = 2 if nelz is None else 3
    nel = nelx * nely * (nelz or 1)
    nn = (nelx + 1) * (nely + 1) * ((nelz or 0) + 1)
    rng = np.random.default_rng(0)
    xPhys = rng.random((nel, 1))
    u = rng.standard_normal((ndim * nn, 2))
    f = rng.standard_normal((ndim * nn, 2))
    export_vtk(str(tmp_path / "d"), nelx, nely, xPhys, nelz, x=xPhys * 0.5, u=u, f=f,
               volfrac=0.4, binary=binary)
    r = read_vtk(tmp_path / "d.vtk")
    assert r["points"].shape == (nn, 3)
    assert np.all(r["types"] == (9 if ndim == 2 else 12))
    P = r["points"][:, :ndim]
    for e, c in enumerate(r["cells"]):
        corners = P[c]
        assert np.allclose(corners.max(0) - corners.min(0), 1.0)  # a unit cell
        if ndim == 2:  # counter-clockwise with y up (positive area)
            xs, ys = corners[:, 0], corners[:, 1]
            assert 0.5 * np.sum(xs * np.roll(ys, -1) - np.roll(xs, -1) * ys) == pytest.approx(1.0)
        else:  # bottom face then top face, same orientation, top at z + 1
            assert np.allclose(corners[4:, 2] - corners[:4, 2], 1.0)
            assert np.allclose(corners[4:, :2], corners[:4, :2])
    # element e of the reference's numbering covers nodes around its centre
    e = nel - 1
    assert np.allclose(P[r["cells"][e]].mean(0)[:2], [nelx - 0.5, 0.5])
    assert np.allclose(r["cell_data"]["xPhys"][:, 0], xPhys[:, 0])
    assert np.allclose(r["cell_data"]["x"][:, 0], 0.5 * xPhys[:, 0])
    assert r["cell_data"]["xThresh"].sum() == np.floor(0.4 * nel)
    for i in range(2):
        assert np.allclose(r["point_data"][f"u{i}"], u[:, i].reshape(nn, ndim))
        assert np.allclose(r["point_data"][f"f{i}"], f[:, i].reshape(nn, ndim))


def test_threshold_keeps_largest():
    x = np.array([[0.1], [0.9], [0.5], [0.7]])
    t = threshold(x, 0.5)
    assert t[:, 0].tolist() == [0.0, 1.0, 0.0, 1.0]
