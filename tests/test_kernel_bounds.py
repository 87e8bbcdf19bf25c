"""CPU restatement of the index ranges that the 3D K.u kernel (ku3d_kernel
in topoptlab_amd/csrc/tl_core.hip) reads from the zero-bordered SIMP array:
every read must stay inside the (nelx+2)(nely+2)(nelz+2) allocation,
including the prefetch of the next element layer on the last node plane
of each z chunk (that layer is clamped to the border layer nelz)."""
import itertools

import pytest

BS = 256


def dens_reads(nelx, nely, nelz):
    nx, ny, nz = nelx + 1, nely + 1, nelz + 1
    nbx = -(-nx * ny // BS)
    zc = nz
    while zc > 6 and nbx * (-(-nz // zc)) < 1024:
        zc = (zc + 1) // 2
    for by in range(-(-nz // zc)):
        z0, z1 = by * zc, min(nz, by * zc + zc)
        # prologue: layers z0-1 and z0; each z step: layer iz+1 (clamped)
        for ez in [z0 - 1, z0] + [min(iz + 1, nelz) for iz in range(z0, z1)]:
            for ix, iy in itertools.product((0, nx - 1), (0, ny - 1)):
                for ox, oy in itertools.product((-1, 0), (-1, 0)):
                    yield ((ez + 1) * (nelx + 2) + ix + ox + 1) * (nely + 2) + iy + oy + 1


@pytest.mark.parametrize("dims", [(32, 16, 16), (16, 8, 8), (24, 16, 20), (192, 96, 96),
                                  (6, 4, 5), (5, 4, 3), (96, 48, 48), (64, 8, 40)])
def test_ku3d_density_reads_in_bounds(dims):
    npad = (dims[0] + 2) * (dims[1] + 2) * (dims[2] + 2)
    idx = list(dens_reads(*dims))
    assert min(idx) >= 0 and max(idx) == npad - 1
