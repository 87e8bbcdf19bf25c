"""CPU restatement of the index ranges of the 3D K.u kernel (ku3d_kernel in
topoptlab_amd/csrc/tl_core.hip) and its launcher:

- the grid (one workgroup per resident slot, at least one per tile column)
  splits the ncol*nz (tile column, node plane) units into one contiguous
  range per workgroup; every unit is walked exactly once, and a range spans
  at most two columns (the two segments the kernel walks);
- every read of the zero-bordered SIMP array stays inside its
  (nelx+2)(nely+2)(nelz+2) allocation, including the prefetch of the next
  element layer on the last node plane of a segment (clamped to the border
  layer nelz);
- every LDS DMA copy of a node plane reads inside the DOF array (offsets are
  clamped) and only clamps entries of nodes outside the grid, and every copy
  writes inside its LDS plane buffer.
"""
import itertools

import pytest

BS = 256


def launch(nelx, nely, nelz, slots):
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    ncol = -(-nx * ny // BS)
    units = ncol * nz
    nwg = max(ncol, min(slots, -(-units // 4)))
    assert units * (nwg + 1) < 2 ** 32
    return ncol, units, nwg


def xcd_order(nb):
    q, r = nb >> 3, nb & 7
    return [(hw & 7) * q + min(hw & 7, r) + (hw >> 3) for hw in range(nb)]


def segments(nelx, nely, nelz, slots):
    """(workgroup, bx, z0, z1) of every segment, in the kernel's order."""
    nz = nelz + 1
    ncol, units, nb = launch(nelx, nely, nelz, slots)
    for hw, w in enumerate(xcd_order(nb)):
        u0, u1 = units * w // nb, units * (w + 1) // nb
        bx = u0 // nz
        z0 = u0 - bx * nz
        z1 = min(nz, z0 + (u1 - u0))
        yield hw, bx, z0, z1
        rest = (u1 - u0) - (z1 - z0)
        if rest > 0:
            yield hw, bx + 1, 0, rest


# the last five: one element, and grids one element thick in z, y or x
TINY = [(1, 1, 1), (1, 1, 3), (3, 2, 1), (2, 1, 2), (1, 3, 2)]
DIMS = [(32, 16, 16), (16, 8, 8), (24, 16, 20), (192, 96, 96), (6, 4, 5), (5, 4, 3),
        (96, 48, 48), (64, 8, 40), (8, 200, 6)] + TINY
SLOTS = [512, 768, 1024, 7, 100]


@pytest.mark.parametrize("slots", SLOTS)
@pytest.mark.parametrize("dims", DIMS)
def test_ku3d_units_walked_once(dims, slots):
    nb = launch(*dims, slots)[2]
    assert sorted(xcd_order(nb)) == list(range(nb))
    ncol, units, _ = launch(*dims, slots)
    nz = dims[2] + 1
    seen = [0] * units
    per_wg = {}
    for hw, bx, z0, z1 in segments(*dims, slots):
        assert 0 <= bx < ncol and 0 <= z0 < z1 <= nz
        per_wg[hw] = per_wg.get(hw, 0) + 1
        for z in range(z0, z1):
            seen[bx * nz + z] += 1
    assert seen == [1] * units
    assert max(per_wg.values()) <= 2 and len(per_wg) == nb


def dens_reads(nelx, nely, nelz, slots):
    nx, ny = nelx + 1, nely + 1
    for _, bx, z0, z1 in segments(nelx, nely, nelz, slots):
        # prologue: layers z0-1 and z0; each z step: layer iz+1 (clamped)
        for ez in [z0 - 1, z0] + [min(iz + 1, nelz) for iz in range(z0, z1)]:
            for ix, iy in itertools.product((0, nx - 1), (0, ny - 1)):
                for ox, oy in itertools.product((-1, 0), (-1, 0)):
                    yield ((ez + 1) * (nelx + 2) + ix + ox + 1) * (nely + 2) + iy + oy + 1


@pytest.mark.parametrize("dims", DIMS)
def test_ku3d_density_reads_in_bounds(dims):
    npad = (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2)
    idx = list(dens_reads(*dims, 512))
    assert min(idx) >= 0 and max(idx) == npad - 1


def plane_dma(ny, ndpn, size):
    """Bytes per lane, copies per plane and LDS bytes per plane buffer."""
    ub = 4
    nu = -(-(BS + 2 * ny + 2) * ndpn * size // ub)
    return ub, nu, -(-nu // BS) * BS * ub


@pytest.mark.parametrize("ndpn,size", [(3, 8), (3, 4), (1, 8), (1, 4)])
@pytest.mark.parametrize("dims", [(32, 16, 16), (6, 4, 5), (24, 16, 20), (8, 200, 6)] + TINY)
def test_ku3d_plane_copies_in_bounds(dims, ndpn, size):
    nelx, nely, nelz = dims
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    plane = nx * ny
    total = plane * nz * ndpn * size
    ub, nu, buf = plane_dma(ny, ndpn, size)
    assert (BS + 2 * ny + 2) * ndpn * size == nu * ub  # whole entries per plane
    for _, bx, z0, z1 in segments(*dims, 512):
        j0 = bx * BS
        # planes copied: z0-1, z0, z0+1 (prologue), iz+2 while a later step needs it
        planes = [z0 - 1, z0, z0 + 1] + [iz + 2 for iz in range(z0, z1) if iz + 1 < z1]
        assert min(planes) >= -1 and max(planes) <= nz
        for pz in planes:
            base = (pz * plane + j0 - ny - 1) * ndpn * size
            for k in range(nu):
                o = base + k * ub
                oc = min(max(o, 0), total - ub)
                assert 0 <= oc and oc + ub <= total
                if oc != o:
                    # a clamped copy only holds entries of nodes outside the grid
                    for b in range(o, o + ub, size):
                        node = b // (ndpn * size)
                        assert node < 0 or node >= plane * nz
                # lane k of round k // BS writes bytes [k*ub, (k+1)*ub) of the buffer
                assert (k + 1) * ub <= buf
