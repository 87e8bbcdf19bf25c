"""Synthetic inputs of the benchmark workloads (host side, numpy).

The 3D cantilever benchmark uses a seeded Gaussian random density field:
white noise on the element grid, Gaussian smoothing with a length scale of
``sigma`` elements (periodic FFT convolution), standardisation, then
x = lo + (1-lo)*Phi(g + b) with the shift b bisected so that mean(x) equals
the requested mean.  Element order is the mesh's own: element id
(iz*nelx + ix)*nely + iy in 3D, nely*ix + iy in 2D.
"""
from __future__ import annotations

import hashlib

import numpy as np
from scipy.special import ndtr


def gaussian_random_density(nelx, nely, nelz=None, seed=0, sigma=4.0,
                            mean=0.3, lo=0.001):
    """Return a float64 element density vector in [lo, 1] with the given mean."""
    rng = np.random.default_rng(seed)
    shape = (nelx, nely) if nelz is None else (nelz, nelx, nely)
    g = rng.standard_normal(shape)
    spec = np.fft.rfftn(g)
    k2 = np.zeros(spec.shape)
    for ax, n in enumerate(shape):
        k = np.fft.fftfreq(n) if ax < len(shape) - 1 else np.fft.rfftfreq(n)
        sh = [1] * len(shape)
        sh[ax] = -1
        k2 = k2 + (k.reshape(sh)) ** 2
    g = np.fft.irfftn(spec * np.exp(-2.0 * np.pi ** 2 * sigma ** 2 * k2), s=shape,
                      axes=tuple(range(len(shape))))
    g = (g - g.mean()) / g.std()
    a, b = -10.0, 10.0
    for _ in range(200):
        m = 0.5 * (a + b)
        if (lo + (1 - lo) * ndtr(g + m)).mean() > mean:
            b = m
        else:
            a = m
    x = lo + (1 - lo) * ndtr(g + 0.5 * (a + b))
    return np.ascontiguousarray(x.reshape(-1), dtype=np.float64)


def field_sha256(x: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(x, dtype=np.float64).tobytes()).hexdigest()


def seeded_loads_2d(rng, nelx, nely, ncases):
    """Unit downward point loads at seeded boundary nodes that are not on the
    clamped left edge (x = 0): nodes on the right, top or bottom edges."""
    nodes = []
    cand = ([(nelx, iy) for iy in range(nely + 1)]
            + [(ix, 0) for ix in range(1, nelx)]
            + [(ix, nely) for ix in range(1, nelx)])
    pick = rng.choice(len(cand), size=ncases, replace=False)
    for p in pick:
        ix, iy = cand[p]
        nodes.append((nely + 1) * ix + iy)
    return np.array(nodes, dtype=np.int64)


def hierarchy_halvings(nelx, nely, ndpn, max_coarse_dof=3000):
    """Fewest grid halvings (at least one) after which a 2D grid of about
    nelx x nely elements, each count rounded up to a multiple of 2^h, has a
    coarsest grid of at most ``max_coarse_dof`` DOFs."""
    h = 1
    while True:
        m = 1 << h
        cx, cy = -(-nelx // m), -(-nely // m)
        if ndpn * (cx + 1) * (cy + 1) <= max_coarse_dof or min(cx, cy) < 2:
            return h
        h += 1


def batched_instances(seed=0, n_elast=32, n_heat=8, xrange=(128, 1024), yrange=(64, 512),
                      ncases=4, max_coarse_dof=3000):
    """Instances of the batched uneven-grid workload: ``n_elast`` 2D Q4
    elasticity grids (left edge clamped, ``ncases`` unit downward point loads
    at seeded boundary nodes) and ``n_heat`` 2D heat-conduction grids (uniform
    source, T = 0 on the left edge).  Element counts are drawn uniformly from
    the ranges; each instance's counts are then rounded up to the multiple of
    2^h that its own multigrid hierarchy needs (h halvings down to a coarsest
    grid of at most ``max_coarse_dof`` DOFs), so a grid is never smaller
    than its draw and at most 2^h - 1 elements larger.  Returns dicts with
    kind, nelx, nely, drawn (the drawn counts), density_seed and
    (elasticity) load_nodes."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_elast + n_heat):
        kind = "elast" if i < n_elast else "heat"
        dx = int(rng.integers(xrange[0], xrange[1] + 1))
        dy = int(rng.integers(yrange[0], yrange[1] + 1))
        m = 1 << hierarchy_halvings(dx, dy, 2 if kind == "elast" else 1, max_coarse_dof)
        nelx, nely = -(-dx // m) * m, -(-dy // m) * m
        d = {"kind": kind, "nelx": nelx, "nely": nely, "drawn": (dx, dy),
             "density_seed": int(rng.integers(0, 2 ** 31))}
        if d["kind"] == "elast":
            d["load_nodes"] = seeded_loads_2d(rng, nelx, nely, ncases)
        out.append(d)
    return out
