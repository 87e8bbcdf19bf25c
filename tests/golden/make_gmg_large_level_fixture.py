"""Write tests/golden/gmg_elast_128x64x64_seed5.npz: the oracle's GMG-PCG
solution of a 3D elasticity problem whose first coarse level is large enough
(65 x 33 x 33 = 70,785 nodes, 3 DOFs each) for the device's large-level
stencil form.

Problem: 128x64x64 H8 elements, x = 0 face clamped in all components,
Gaussian random density (seed 5, sigma 2), SIMP stiffness, right-hand side
standard normal on the free DOFs (numpy default_rng(3)).  Solver: the oracle's
PCG (linear_solvers.py:50-136) preconditioned by its FP64 V(1,1)-cycle over 4
levels (multigrid.py:11-223, Galerkin coarse operators, damped Jacobi 0.6,
exact coarsest solve), relative residual 1e-8.

The fixture keeps the iteration count, b.u, ||u||_2 and u at 4096 seeded
free DOFs (with their indices).  The oracle takes about 90 s on one core, too
long for a GPU test, so it runs once here and the result is committed.
Usage: python tests/golden/make_gmg_large_level_fixture.py
"""
from __future__ import annotations

import hashlib
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle import fem, mg  # noqa: E402
from topoptlab_amd.workloads import gaussian_random_density  # noqa: E402

DIMS = (128, 64, 64)
NLEVELS = 4
OUT = Path(__file__).resolve().parent / "gmg_elast_128x64x64_seed5.npz"


def problem():
    """K_ff, the free DOFs and the density of the fixture's problem."""
    nelx, nely, nelz = DIMS
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    iz, iy = np.meshgrid(np.arange(nelz + 1), np.arange(nely + 1), indexing="ij")
    nodes = (iz * (nelx + 1) * (nely + 1) + iy).ravel()
    fixed = np.sort((3 * nodes[:, None] + np.arange(3)[None, :]).ravel())
    free = np.setdiff1d(np.arange(ndof), fixed)
    x = gaussian_random_density(nelx, nely, nelz, seed=5, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(fem.ke_elast_3d(), fem.edofmat_3d(nelx, nely, nelz, 3),
                                    fem.simp(x), ndof), free).tocsr()
    return K, fixed, free, x


def main():
    t = time.perf_counter()
    K, fixed, free, x = problem()
    b = np.random.default_rng(3).standard_normal(len(free))
    h = mg.Hierarchy(K, *DIMS, 3, fixed, NLEVELS)
    u, info, its = mg.pcg(K, b, h.vcycle, rtol=1e-8)
    assert info == 0
    relres = np.linalg.norm(K @ u - b) / np.linalg.norm(b)
    print(f"oracle GMG-PCG {time.perf_counter() - t:.1f} s, {its} iterations, relres {relres:.2e}")
    idx = np.sort(np.random.default_rng(0).choice(len(free), 4096, replace=False))
    np.savez(OUT, dims=np.array(DIMS), nlevels=NLEVELS, iterations=its, relres=relres,
             bu=float(b @ u), unorm=float(np.linalg.norm(u)), idx_free=idx, u_idx=u[idx],
             density_sha256=hashlib.sha256(x.tobytes()).hexdigest())
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
