"""Write tests/golden/c2_oracle_192x96x96.npz: the oracle's FP64 GMG-PCG
solution of the 192x96x96 H8 cantilever (the headline workload: x = 0 face
clamped, unit line load along the bottom edge of the x = L face, Gaussian
random field density seed 0, SIMP p = 3, Emin 1e-9).

Solver: the oracle's PCG (linear_solvers.py:50-136) preconditioned by its
FP64 V(1,1)-cycle over 5 levels (multigrid.py:11-223: Galerkin coarse
operators, damped Jacobi 0.6, direct coarsest solve), relative residual 1e-8.

The fixture keeps the iteration count, the true relative residual, the
compliance f.u, ||u||_2 and u at 4096 seeded free DOFs (with their indices).
The oracle needs about 75 s and 47 GiB of host memory (431M nonzeros), so it
runs once and the result is committed.
Usage: python tests/golden/make_c2_oracle_fixture.py
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

DIMS = (192, 96, 96)
NLEVELS = 5
OUT = Path(__file__).resolve().parent / "c2_oracle_192x96x96.npz"


def main():
    nelx, nely, nelz = DIMS
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    x = gaussian_random_density(nelx, nely, nelz, seed=0)
    f, fixed, free = fem.clamped_lineload_3d(nelx, nely, nelz, ndof)
    t = time.perf_counter()
    K = fem.apply_bc(fem.assemble_K(fem.ke_elast_3d(), fem.edofmat_3d(nelx, nely, nelz, 3),
                                    fem.simp(x), ndof), free).tocsr()
    print(f"assembled in {time.perf_counter() - t:.1f} s", flush=True)
    h = mg.Hierarchy(K, nelx, nely, nelz, 3, fixed, NLEVELS)
    print(f"hierarchy in {time.perf_counter() - t:.1f} s", flush=True)
    b = f[free, 0]
    u, info, its = mg.pcg(K, b, h.vcycle, rtol=1e-8)
    assert info == 0
    relres = float(np.linalg.norm(K @ u - b) / np.linalg.norm(b))
    print(f"oracle GMG-PCG {time.perf_counter() - t:.1f} s, {its} iterations, relres {relres:.3e}",
          flush=True)
    idx = np.sort(np.random.default_rng(0).choice(len(free), 4096, replace=False))
    np.savez(OUT, dims=np.array(DIMS), nlevels=NLEVELS, iterations=its, relres=relres,
             compliance=float(b @ u), unorm=float(np.linalg.norm(u)), idx_free=idx,
             u_idx=u[idx], density_sha256=hashlib.sha256(x.tobytes()).hexdigest())
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
