"""Host-side modes of bench.py that need no GPU: the refusal of several GPUs
and the full-size CPU timing mode (run here on a small grid and checked
against a direct solve of the same system)."""
import json
import subprocess
import sys
from pathlib import Path

import numpy as np
from scipy.sparse.linalg import spsolve

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def run_bench(*args):
    return subprocess.run([sys.executable, str(ROOT / "bench.py"), *args], capture_output=True,
                          text=True, timeout=300, cwd=ROOT)


def test_several_gpus_are_refused():
    r = run_bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_full_mode_small_grid():
    r = run_bench("--cpu-full", "--nelx", "16", "--levels", "3")
    assert r.returncode == 0, r.stderr
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cpu_full"] == "16x8x8" and out["levels"] == 3 and out["converged"]
    assert out["relres_true"] < 1e-8 and out["pcg_iterations"] > 0
    from oracle import fem
    from topoptlab_amd.workloads import gaussian_random_density
    nelx, nely, nelz = 16, 8, 8
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    x = gaussian_random_density(nelx, nely, nelz, seed=0)
    f, fixed, free = fem.clamped_lineload_3d(nelx, nely, nelz, ndof)
    K = fem.apply_bc(fem.assemble_K(fem.ke_elast_3d(), fem.edofmat_3d(nelx, nely, nelz, 3),
                                    fem.simp(x), ndof), free).tocsc()
    b = f[free, 0]
    c = float(b @ spsolve(K, b))
    assert abs(out["compliance"] - c) / abs(c) < 1e-7
