"""CPU restatement of the index ranges of the grouped 3D stencil form
(st_apply_kernel with st_row_3d in topoptlab_amd/csrc/tl_mg.hip):

- the XCD-aware block order (logical = xcd*q + min(xcd, r) + hw/8 over nb
  blocks) is a permutation, so with 64-thread blocks every node is one
  thread's row exactly once;
- every raw buffer load reads inside its descriptor: x (nnodes*NDPN values)
  and each forward-half slab kk (NDPN^2 arrays of nnodes values), with the
  32-bit byte offsets the kernel forms (voffset + soffset) below the slab
  size, for every one of the 27 terms and every node, including clamped
  neighbours outside the grid, which carry weight 0;
- the host's size rule (nnodes*NDPN^2*8 < 2^32) keeps those offsets inside
  32 bits at the largest admissible level.
"""
import numpy as np
import pytest


def logical_blocks(nb):
    hw = np.arange(nb)
    xcd, q, r = hw & 7, nb >> 3, nb & 7
    return xcd * q + np.minimum(xcd, r) + (hw >> 3)


@pytest.mark.parametrize("nb", [1, 7, 8, 9, 120, 479, 3639, 4096 + 5])
def test_block_order_is_a_permutation(nb):
    lg = logical_blocks(nb)
    assert np.array_equal(np.sort(lg), np.arange(nb))


def term_offsets(nx, ny, nz, ndpn, k):
    """Byte offsets (x loads, block loads) of term k at every node, as the
    kernel forms them, and the neighbour-inside flag."""
    nnodes = nx * ny * nz
    n = np.arange(nnodes, dtype=np.int64)
    iy, t = n % ny, n // ny
    ix, iz = t % nx, t // nx
    rz, rx, ry = k // 9 - 1, k // 3 % 3 - 1, k % 3 - 1
    m = (np.clip(iz + rz, 0, nz - 1) * nx + np.clip(ix + rx, 0, nx - 1)) * ny + \
        np.clip(iy + ry, 0, ny - 1)
    ok = (ix + rx >= 0) & (ix + rx < nx) & (iy + ry >= 0) & (iy + ry < ny) & \
         (iz + rz >= 0) & (iz + rz < nz)
    arr = nnodes * 8
    xo = [m * ndpn * 8 + c * 8 for c in range(ndpn)]
    if k >= 13:
        bo = [n * 8 + e * arr for e in range(ndpn * ndpn)]
    else:
        bo = [m * 8 + (c * ndpn + d) * arr for d in range(ndpn) for c in range(ndpn)]
    return xo, bo, ok, m


@pytest.mark.parametrize("dims", [(2, 2, 2), (5, 3, 4), (9, 5, 5), (3, 17, 2), (33, 17, 17),
                                  (13, 7, 3), (7, 4, 2)])
@pytest.mark.parametrize("ndpn", [1, 3])
def test_group_form_buffer_loads_in_bounds(dims, ndpn):
    nx, ny, nz = dims
    nnodes = nx * ny * nz
    x_bytes = nnodes * ndpn * 8
    slab_bytes = nnodes * ndpn * ndpn * 8
    seen_kk = set()
    for k in range(27):
        xo, bo, ok, m = term_offsets(nx, ny, nz, ndpn, k)
        for o in xo:
            assert o.min() >= 0 and o.max() + 8 <= x_bytes
        for o in bo:
            assert o.min() >= 0 and o.max() + 8 <= slab_bytes
        kk = k - 13 if k >= 13 else 13 - k
        assert 0 <= kk <= 13
        seen_kk.add(kk)
        # neighbours outside the grid are clamped onto it (weight 0)
        assert np.all((m >= 0) & (m < nnodes))
        if k == 13:
            assert ok.all()
    assert seen_kk == set(range(14))


def test_size_rule_keeps_offsets_in_32_bits():
    for ndpn in (1, 3):
        # largest node count the host admits for the grouped form
        nmax = ((1 << 32) - 1) // (ndpn * ndpn * 8)
        assert nmax * ndpn * ndpn * 8 < 1 << 32
        # the largest voffset + soffset a load forms stays below the slab size
        assert (nmax - 1) * 8 + (ndpn * ndpn - 1) * nmax * 8 < 1 << 32
        assert (nmax - 1) * ndpn * 8 + (ndpn - 1) * 8 < 1 << 32


@pytest.mark.parametrize("dims", [(2, 2, 2), (5, 3, 4), (13, 7, 7), (25, 13, 13), (3, 17, 2)])
@pytest.mark.parametrize("ndpn", [1, 3])
def test_lane_form_covers_each_term_once_in_bounds(dims, ndpn):
    """st_apply_quad_kernel (four lanes per node, 64-thread workgroups, the
    XCD-aware block order; lane q sums the terms K = q, q + 4, ... and writes
    DOF q): the live terms of a node's lanes are its 27 terms exactly once,
    every DOF is written by exactly one lane, and every read (x of the
    clamped neighbour, the forward-half block) is inside its buffer."""
    lanes = 4
    nx, ny, nz = dims
    nnodes = nx * ny * nz
    nn, nt = ndpn * ndpn, 7
    ng = -(-27 // (lanes * nt))
    nb = -(-lanes * nnodes // 64)
    t = logical_blocks(nb)[:, None] * 64 + np.arange(64)[None, :]
    n, q = (t // lanes).ravel(), (t % lanes).ravel()
    keep = n < nnodes
    n, q = n[keep], q[keep]
    iy, r = n % ny, n // ny
    ix, iz = r % nx, r // nx
    terms = np.zeros((nnodes, 27), dtype=np.int64)
    for j in range(ng * nt):
        kraw = q + lanes * j
        live = kraw <= 26
        k = np.minimum(kraw, 26)
        rz, rx, ry = k // 9 - 1, k // 3 % 3 - 1, k % 3 - 1
        m = (np.clip(iz + rz, 0, nz - 1) * nx + np.clip(ix + rx, 0, nx - 1)) * ny + \
            np.clip(iy + ry, 0, ny - 1)
        assert m.min() >= 0 and (m * ndpn + ndpn - 1).max() < nnodes * ndpn
        fwd = k >= 13
        slab = np.where(fwd, k - 13, 13 - k)
        base = np.where(fwd, n, m)
        idx = slab * nn * nnodes + (nn - 1) * nnodes + base
        assert idx.min() >= 0 and idx.max() < 14 * nn * nnodes
        np.add.at(terms, (n[live], k[live]), 1)
    assert (terms == 1).all()
    writes = np.zeros((nnodes, ndpn), dtype=np.int64)
    for d in range(ndpn):
        sel = d % lanes == q
        np.add.at(writes, (n[sel], d), 1)
    assert (writes == 1).all()
