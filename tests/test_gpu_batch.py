"""GPU parity of the batched uneven-grid solver (marked ``gpu``).

* Every 2D operator workgroup shape gives K.u equal to the assembled oracle
  matrix to 1e-13 and bit-identical to the node-line shape (same sums in the
  same order), on grids whose sizes do not divide the tiles.
* The autotuner returns a valid shape and reuses its table.
* ``batch_solve`` (hipGraph-replayed PCG chunks on several streams) gives
  the same iteration counts and bit-identical solutions as one-at-a-time
  ``solve`` calls, with graphs on, off, and on again (cached graphs
  replayed), and its solutions meet the tolerance against the oracle's
  assembled matrix; also for a batch of one instance with a 1-D right-hand
  side and a batch mixing 2D and 3D instances.
"""
import sys

import numpy as np
import pytest
from scipy.sparse.linalg import spsolve

from oracle import fem
from tests.conftest import rel

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _case(nelx, nely, kind, seed):
    from topoptlab_amd.workloads import gaussian_random_density
    if kind == "elast":
        KE, ndpn, eps = fem.ke_elast_2d(), 2, 1e-9
    else:
        KE, ndpn, eps = fem.ke_poisson_2d(), 1, 1e-3
    edof = fem.edofmat_2d(nelx, nely, ndpn)
    ndof = ndpn * (nelx + 1) * (nely + 1)
    fixed = np.arange(ndpn * (nely + 1))  # left edge
    free = np.setdiff1d(np.arange(ndof), fixed)
    x = gaussian_random_density(nelx, nely, seed=seed, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x, eps), ndof), free).tocsr()
    return KE, eps, ndof, fixed, free, x, K


@pytest.mark.parametrize("nelx,nely,kind", [(37, 21, "elast"), (40, 24, "heat"),
                                            (8, 70, "elast"), (129, 5, "heat")])
def test_tile_shapes_match_assembled(gpu, nelx, nely, kind):
    import torch
    from topoptlab_amd.batch import TILES, set_tile
    from topoptlab_amd.fem import StiffnessOperator
    say("tiles", nelx, nely, kind)
    KE, eps, ndof, fixed, free, x, K = _case(nelx, nely, kind, seed=nelx)
    op = StiffnessOperator(nelx, nely, KE=KE, fixed=fixed, eps=eps, device=gpu)
    op.set_density(x)
    u = np.random.default_rng(1).standard_normal(len(free))
    ut = torch.as_tensor(u, device=gpu)
    y_ref = K @ u
    outs = {}
    for ty in TILES:
        set_tile(op, "fp64", ty)
        set_tile(op, "fp32", ty)
        y = (op @ ut).cpu().numpy()
        assert rel(y, y_ref) < 1e-13, ty
        outs[ty] = y
        y32 = op.matvec_full(op.full(ut), "fp32")[op._free_t].cpu().numpy()
        assert rel(y32, y_ref) < 1e-5, ty
    for ty in TILES[1:]:
        assert np.array_equal(outs[ty], outs[0]), ty
    # The fused PCG product (p = z + beta p; q = K p; p.q) through a solve.
    # Each shape partitions p.q into other blocks, so the sums round
    # differently: on a moderate-contrast density the solves agree closely.
    op.set_density(np.full(nelx * nely, 0.5))
    f = op.full(torch.as_tensor(np.random.default_rng(2).standard_normal(len(free)), device=gpu))
    sols = {}
    for ty in TILES:
        set_tile(op, "fp64", ty)
        s, inf = op.solve(f, rtol=1e-8, preconditioner="jacobi", maxiter=100000)
        assert inf["relres_true"] < 1.1e-8, ty
        sols[ty] = (s.cpu().numpy(), inf["iterations"])
    for ty in TILES[1:]:
        assert abs(sols[ty][1] - sols[0][1]) <= 1, (ty, sols[ty][1], sols[0][1])
        assert rel(sols[ty][0], sols[0][0]) < 1e-6, ty
    op.close()


def test_autotune_table(gpu):
    from topoptlab_amd.batch import TILES, autotune, clear_autotune_table
    from topoptlab_amd.elements import lk_linear_elast_2d, lk_linear_elast_3d
    from topoptlab_amd.fem import StiffnessOperator
    from topoptlab_amd.workloads import gaussian_random_density
    say("autotune")
    clear_autotune_table()
    ops = []
    for _ in range(2):
        op = StiffnessOperator(256, 96, KE=lk_linear_elast_2d(), fixed=np.arange(194), device=gpu)
        op.set_density(gaussian_random_density(256, 96, seed=0))
        ops.append(op)
    t = autotune(ops[0], reps=5)
    assert t["fp64"] in TILES and t["fp32"] in TILES
    assert all(v > 0 for v in t["us_fp64"].values()) and all(v > 0 for v in t["us_fp32"].values())
    assert t["us_fp64"][t["fp64"]] == min(t["us_fp64"].values())
    assert autotune(ops[1], reps=5) == t  # same grid: the table's entry
    op3 = StiffnessOperator(4, 4, 4, KE=lk_linear_elast_3d(), device=gpu)
    op3.set_density(np.full(64, 0.5))
    assert autotune(op3)["fp64"] == 0
    for op in ops + [op3]:
        op.close()


def test_batch_matches_single_solves(gpu):
    import torch
    from topoptlab_amd.batch import autotune, batch_setup, batch_solve, build_instance
    from topoptlab_amd.workloads import batched_instances
    specs = batched_instances(seed=3, n_elast=4, n_heat=2, xrange=(64, 192), yrange=(32, 128),
                              ncases=3)
    ops, fs, xs, lvs = [], [], [], []
    for spec in specs:
        say("batch instance", spec["kind"], spec["nelx"], spec["nely"])
        op, x, f, levels = build_instance(spec, gpu)
        op.set_density(x)
        op.setup_mg(nlevels=levels, precision="fp32")
        autotune(op, reps=3)
        ops.append(op)
        xs.append(x)
        lvs.append(levels)
        fs.append(torch.as_tensor(f, device=gpu))
    ref = [[op.solve(f[k], rtol=1e-8) for k in range(f.shape[0])] for op, f in zip(ops, fs)]
    # the set-ups again, overlapped on 3 streams: the batch solves below must
    # still equal the single solves after the one-after-another set-up
    say("batch set-up on 3 streams")
    batch_setup(ops, [torch.as_tensor(x, device=gpu) for x in xs], lvs, precision="fp32",
                nstreams=3)
    # lockstep (one launch per kernel and place in the cycle for all systems;
    # chunks captured as graphs and replayed, and eager), then per-instance
    # streams with and without captured graphs; the graph forms twice in a row
    # (the second call reuses what the first set up)
    for lockstep, graphs in ((True, True), (True, True), (True, False), (False, True),
                             (False, False), (False, True)):
        say("batch solve, lockstep", lockstep, "graphs", graphs)
        U, infos = batch_solve(ops, fs, rtol=1e-8, nstreams=3, chunk=4, use_graphs=graphs,
                               lockstep=lockstep)
        for i in range(len(ops)):
            assert U[i].shape == fs[i].shape
            for k in range(fs[i].shape[0]):
                u_ref, inf_ref = ref[i][k]
                assert infos[i][k]["iterations"] == inf_ref["iterations"], (i, k, lockstep, graphs)
                assert torch.equal(U[i][k], u_ref), (i, k, lockstep, graphs)
                # PCG stops on the recursive residual (linear_solvers.py:107-117).
                # A point load on near-void material (Emin 1e-9) gives |u| up
                # to ~1e7 and the recomputed residual sits above the recursive
                # one: the oracle's FP64 GMG-PCG on these instances ends at
                # true relres up to 8.3e-8, the device at 7.7e-8.
                assert infos[i][k]["relres_recursive"] < 1e-8
                assert infos[i][k]["relres_true"] < 1e-7
    # against the oracle's assembled matrix: the two smallest instances
    order = sorted(range(len(specs)), key=lambda i: ops[i].ndof)[:2]
    for i in order:
        spec = specs[i]
        kind = spec["kind"]
        KE, eps, ndof, fixed, free, _, _ = _case(spec["nelx"], spec["nely"], kind, 0)
        ndpn = 2 if kind == "elast" else 1
        edof = fem.edofmat_2d(spec["nelx"], spec["nely"], ndpn)
        K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(xs[i], eps), ndof), free).tocsc()
        for k in range(fs[i].shape[0]):
            b = fs[i][k].cpu().numpy()[free]
            u = U[i][k].cpu().numpy()[free]
            assert np.linalg.norm(b - K @ u) / np.linalg.norm(b) < 1e-7, (i, k)
            assert rel(u, spsolve(K, b)) < 1e-4, (i, k)
    for op in ops:
        op.close()


def test_batch_reports_non_convergence(gpu):
    """maxiter too small: TL_ENOCONV from the batch, every system's status
    says so, and raise_on_fail=False returns the partial results."""
    import torch
    from topoptlab_amd.batch import batch_solve, build_instance
    from topoptlab_amd.workloads import batched_instances
    specs = batched_instances(seed=4, n_elast=2, n_heat=1, xrange=(64, 96), yrange=(32, 64),
                              ncases=2)
    ops, fs = [], []
    for spec in specs:
        op, x, f, levels = build_instance(spec, gpu)
        op.set_density(x)
        op.setup_mg(nlevels=levels, precision="fp32")
        ops.append(op)
        fs.append(torch.as_tensor(f, device=gpu))
    for lockstep in (True, False):
        say("batch non-convergence, lockstep", lockstep)
        with pytest.raises(RuntimeError):
            batch_solve(ops, fs, rtol=1e-12, maxiter=2, chunk=2, lockstep=lockstep)
        U, infos = batch_solve(ops, fs, rtol=1e-12, maxiter=2, chunk=2, raise_on_fail=False,
                               lockstep=lockstep)
        for lst in infos:
            for inf in lst:
                assert inf["status"] != 0 and inf["iterations"] <= 4
    for op in ops:
        op.close()


def test_batch_single_instance_and_mixed_dimensions(gpu):
    """A batch of one instance with a one-dimensional right-hand side, and a
    batch that mixes a 2D and a 3D instance (not eligible for lockstep, so
    the per-instance streams run it even when lockstep is asked for): both
    equal the single solves bit for bit."""
    import torch
    from topoptlab_amd.batch import batch_solve, build_instance
    from topoptlab_amd.fem import StiffnessOperator
    from topoptlab_amd.workloads import batched_instances, gaussian_random_density
    from tests.test_gpu_parity import _problem
    spec = batched_instances(seed=6, n_elast=1, n_heat=0, xrange=(64, 96), yrange=(32, 64),
                             ncases=2)[0]
    op2, x2, f2, lv2 = build_instance(spec, gpu)
    op2.set_density(x2)
    op2.setup_mg(nlevels=lv2, precision="fp32")
    f2 = torch.as_tensor(f2, device=gpu)
    KE, edof, ndof, fixed, free, _ = _problem((8, 4, 4), 3)
    op3 = StiffnessOperator(8, 4, 4, KE=KE, fixed=fixed, device=gpu)
    op3.set_density(gaussian_random_density(8, 4, 4, seed=5, sigma=2.0))
    op3.setup_mg(nlevels=2, precision="fp32")
    f3 = torch.zeros(ndof, dtype=torch.float64, device=gpu)
    f3[torch.as_tensor(free, device=gpu)] = torch.as_tensor(
        np.random.default_rng(7).standard_normal(len(free)), device=gpu)
    ref2 = op2.solve(f2[0], rtol=1e-8)
    ref3 = op3.solve(f3, rtol=1e-8)
    for lockstep in (True, False):
        say("batch of one, 1-D right-hand side, lockstep", lockstep)
        U, infos = batch_solve([op2], [f2[0]], rtol=1e-8, chunk=4, lockstep=lockstep)
        assert U[0].shape == f2[0].shape and len(infos[0]) == 1
        assert infos[0][0]["iterations"] == ref2[1]["iterations"]
        assert torch.equal(U[0], ref2[0])
        say("mixed 2D and 3D batch, lockstep", lockstep)
        U, infos = batch_solve([op2, op3], [f2[0], f3], rtol=1e-8, chunk=4, lockstep=lockstep)
        assert torch.equal(U[0], ref2[0]) and torch.equal(U[1], ref3[0])
        assert infos[1][0]["iterations"] == ref3[1]["iterations"]
    op2.close()
    op3.close()
