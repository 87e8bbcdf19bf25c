"""GPU parity of the HIP path against the CPU oracle (marked ``gpu``).

Tolerances: FP64 kernels 1e-12 relative (matrix-free summation order differs
from scipy's CSR product); filter and OC kernels bit-exact (same summation
order, no FMA contraction); FP32 operator 1e-5; solves compared after both
sides converge to a tight residual.
"""
import sys

import numpy as np
import pytest
from scipy.sparse.linalg import spsolve

from oracle import fem, mg, topopt
from tests.conftest import rel
from topoptlab_amd import elements, example_bc, workloads

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _problem(dims, ndpn, seed=0, bc="cantilever"):
    nelx, nely, nelz = dims
    ndim = 2 if nelz is None else 3
    if ndim == 2:
        KE = fem.ke_elast_2d() if ndpn == 2 else fem.ke_poisson_2d()
        edof = fem.edofmat_2d(nelx, nely, ndpn)
    else:
        KE = fem.ke_elast_3d() if ndpn == 3 else elements.lk_poisson_3d()
        edof = fem.edofmat_3d(nelx, nely, nelz, ndpn)
    nn = (nelx + 1) * (nely + 1) * (1 if nelz is None else nelz + 1)
    ndof = ndpn * nn
    # clamp the x = 0 face/edge (all components)
    if ndim == 2:
        nodes = np.arange(nely + 1)
    else:
        iz, iy = np.meshgrid(np.arange(nelz + 1), np.arange(nely + 1), indexing="ij")
        nodes = (iz * (nelx + 1) * (nely + 1) + iy).ravel()
    fixed = np.sort((ndpn * nodes[:, None] + np.arange(ndpn)[None, :]).ravel())
    if bc == "none":
        fixed = np.array([], dtype=np.int64)
    free = np.setdiff1d(np.arange(ndof), fixed)
    rng = np.random.default_rng(seed)
    x = 0.001 + 0.999 * rng.random(edof.shape[0])
    return KE, edof, ndof, fixed, free, x


# 3D cases cover one tile column (6x4x5), workgroup ranges that cross tile
# columns (24x16x20, 40x30x6) and node planes wider than a workgroup in y, so
# that each plane copy takes several rounds (4x300x3: ny = 301 > 256)
CASES = [((12, 8, None), 2), ((10, 6, None), 1), ((6, 4, 5), 3), ((5, 4, 3), 1),
         ((24, 16, 20), 3), ((300, 7, None), 2), ((4, 300, 3), 3), ((40, 30, 6), 1)]
# one element, and grids one element thick (the plane walk with two node
# planes, node lines of two nodes)
TINY = [((1, 1, None), 2), ((5, 1, None), 1), ((1, 4, None), 2), ((1, 1, 1), 3),
        ((3, 2, 1), 3), ((1, 3, 2), 1), ((2, 1, 2), 3), ((1, 1, 3), 1)]


@pytest.mark.parametrize("dims,ndpn", CASES + TINY)
def test_ku_matches_assembled(gpu, dims, ndpn):
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("K.u", dims, ndpn)
    KE, edof, ndof, fixed, free, x = _problem(dims, ndpn)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    u = np.random.default_rng(1).standard_normal(len(free))
    y_ref = K @ u
    y = (op @ torch.as_tensor(u, device=gpu)).cpu().numpy()
    assert rel(y, y_ref) < 1e-13
    y32 = op.matvec_full(op.full(torch.as_tensor(u, device=gpu)), "fp32")[op._free_t].cpu().numpy()
    assert rel(y32, y_ref) < 1e-5
    d = op.diagonal().cpu().numpy()
    assert rel(d, K.diagonal()) < 1e-14
    # full-length input: fixed entries are ignored, fixed outputs are zero
    uf = np.random.default_rng(2).standard_normal(ndof)
    yf = op.matvec_full(torch.as_tensor(uf, device=gpu)).cpu().numpy()
    assert np.all(yf[fixed] == 0)
    assert rel(yf[free], K @ uf[free]) < 1e-13
    op.close()


@pytest.mark.parametrize("dims,ndpn", [((12, 8, None), 2), ((6, 4, 5), 3), ((10, 6, None), 1)])
def test_ku_sheared_elements_generic_path(gpu, dims, ndpn):
    """Element matrices of non-unit, sheared elements (elements.py with l and
    g) do not match the unit-cell value pattern, so the device takes its
    generic element-matrix path: K.u against the assembled matrix."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("ku sheared", dims, ndpn)
    nelx, nely, nelz = dims
    if nelz is None:
        KE = (elements.lk_linear_elast_2d(l=np.array([1.5, 0.75]), g=np.array([0.3])) if ndpn == 2
              else elements.lk_poisson_2d(l=np.array([1.5, 0.75]), g=np.array([0.3])))
        edof = fem.edofmat_2d(nelx, nely, ndpn)
    else:
        KE = elements.lk_linear_elast_3d(l=np.array([1.2, 0.9, 0.6]), g=np.array([0.2, -0.1]))
        edof = fem.edofmat_3d(nelx, nely, nelz, ndpn)
    _, _, ndof, fixed, free, x = _problem(dims, ndpn)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free).tocsr()
    op = StiffnessOperator(nelx, nely, nelz, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    u = np.random.default_rng(2).standard_normal(len(free))
    y = (op @ torch.as_tensor(u, device=gpu)).cpu().numpy()
    assert rel(y, K @ u) < 1e-13
    op.close()


@pytest.mark.parametrize("dims,ndpn,nlevels,dit", [
    ((16, 8, 8), 3, 3, 1), ((32, 16, None), 2, 4, 1), ((16, 16, None), 1, 3, 1),
    ((8, 4, 4), 3, 2, 1),
    # thin: levels of 3 node planes / lines; the 32x4 strip is so ill-conditioned
    # (true residual 75x the recursive one at the stop, in the oracle too) that
    # rounding moves the iteration count by a few (device 36, oracle 39)
    ((24, 12, 4), 3, 3, 1), ((32, 4, None), 2, 3, 4)])
def test_pcg_gmg_fp64_matches_oracle(gpu, dims, ndpn, nlevels, dit):
    """With an FP64 V-cycle the device preconditioner is the oracle's V(1,1)
    (same Galerkin hierarchy, omega=0.6, zero coarse guesses, exact coarsest
    solve): the first PCG iterates agree to rounding, and PCG takes the same
    number of iterations (within dit)."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("gmg fp64", dims, ndpn, nlevels)
    KE, edof, ndof, fixed, free, _ = _problem(dims, ndpn)
    x = workloads.gaussian_random_density(*dims[:2], dims[2], seed=5, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free).tocsr()
    rng = np.random.default_rng(3)
    b = rng.standard_normal(len(free))
    h = mg.Hierarchy(K, *dims, ndpn, fixed, nlevels)
    u_ref, info, it_ref = mg.pcg(K, b, h.vcycle, rtol=1e-10)
    assert info == 0
    # both stop on the recursive residual (linear_solvers.py:107-117); on the
    # ill-conditioned thin grids the true residual then stays above rtol
    # (oracle: 2.9e-10 on 24x12x4, 7.5e-9 on 32x4), so the device's true
    # residual is held to the oracle's own
    tr_ref = max(2e-10, 2 * np.linalg.norm(K @ u_ref - b) / np.linalg.norm(b))
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=nlevels, precision="fp64")
    # the preconditioner itself: iterates 1 and 3 (x1 = alpha M^-1 b, ...)
    # agree to rounding, which the coarsest solves (dense inverse here, a
    # sparse direct solve in the oracle) and the PCG recurrences amplify by
    # the operators' condition (iterate 3: 3.7e-10 on 8x4x4, 1.1e-8 on the
    # 32x4 strip); mistakes planted in the oracle (smoother weight, a missing
    # sweep, a missing interpolation term) move them by 3.6e-3 or more,
    # test_oracle.py::test_early_pcg_iterates_expose_vcycle_mistakes
    for k, tol in ((1, 1e-8), (3, 1e-7)):
        xk_ref, _, _ = mg.pcg(K, b, h.vcycle, rtol=1e-10, maxiter=k)
        xk, infk = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10, maxiter=k,
                            raise_on_fail=False)
        assert infk["iterations"] == k
        assert rel(xk[op._free_t].cpu().numpy(), xk_ref) < tol, k
    u, inf = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10)
    u = u[op._free_t].cpu().numpy()
    assert abs(inf["iterations"] - it_ref) <= dit, (inf, it_ref)
    assert rel(u, u_ref) < 1e-7
    ud = spsolve(K.tocsc(), b)
    assert rel(u, ud) < 1e-6
    assert inf["relres_true"] < tr_ref
    # FP32 V-cycle converges to the same tolerance in a similar count (up to
    # 30% more on the ill-conditioned thin plate: 65 against 50)
    op.setup_mg(nlevels=nlevels, precision="fp32")
    u32, inf32 = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10)
    assert inf32["relres_true"] < 2 * tr_ref and inf32["iterations"] <= 1.35 * it_ref + 2
    # Jacobi-PCG reaches the same solution
    uj, infj = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10,
                        preconditioner="jacobi", maxiter=200000)
    assert rel(uj[op._free_t].cpu().numpy(), ud) < 1e-6
    op.close()


def test_pcg_gmg_fp64_matches_oracle_large_level(gpu):
    """128x64x64 H8 heat conduction (545k DOF), 4 levels: level 1 has 70,785
    nodes (3,639 workgroups of the grouped 3D stencil form, more than the
    device holds at once) with one DOF per node (the 3-DOF case is the
    fixture test below), against the oracle's V-cycle."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    dims, ndpn, nlevels = (128, 64, 64), 1, 4
    say("gmg fp64 large level", dims)
    KE, edof, ndof, fixed, free, _ = _problem(dims, ndpn)
    x = workloads.gaussian_random_density(*dims, seed=5, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free).tocsr()
    b = np.random.default_rng(3).standard_normal(len(free))
    h = mg.Hierarchy(K, *dims, ndpn, fixed, nlevels)
    u_ref, info, it_ref = mg.pcg(K, b, h.vcycle, rtol=1e-8)
    assert info == 0
    say("oracle iterations", it_ref)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=nlevels, precision="fp64")
    u, inf = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-8)
    u = u[op._free_t].cpu().numpy()
    assert abs(inf["iterations"] - it_ref) <= 1, (inf, it_ref)
    assert rel(u, u_ref) < 1e-6
    assert inf["relres_true"] < 2e-8
    op.close()


def test_pcg_gmg_fp64_large_level_elasticity_fixture(gpu):
    """128x64x64 H8 elasticity, 4 levels: level 1 has 70,785 nodes with 3
    DOFs each, a size class of the 192x96x96 workload's level 1 (232,897
    nodes) that the small cases do not reach.  Against the oracle's FP64
    GMG-PCG solution committed by tests/golden/make_gmg_large_level_fixture.py
    (about 90 s of CPU): the same iteration count (+-1), u at 4096 seeded
    free DOFs to 1e-6, b.u to 1e-7, both at relres 1e-8."""
    import hashlib
    import torch
    from tests.conftest import GOLDEN
    from topoptlab_amd.fem import StiffnessOperator
    fx = np.load(GOLDEN / "gmg_elast_128x64x64_seed5.npz")
    dims = tuple(int(v) for v in fx["dims"])
    say("gmg fp64 large level elasticity fixture", dims)
    KE, _, _, fixed, free, _ = _problem(dims, 3)
    x = workloads.gaussian_random_density(*dims, seed=5, sigma=2.0)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(fx["density_sha256"])
    b = np.random.default_rng(3).standard_normal(len(free))
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=int(fx["nlevels"]), precision="fp64")
    u, inf = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-8)
    u = u[op._free_t].cpu().numpy()
    say("iterations", inf["iterations"], "oracle", int(fx["iterations"]))
    assert abs(inf["iterations"] - int(fx["iterations"])) <= 1
    assert rel(u[fx["idx_free"]], fx["u_idx"]) < 1e-6
    assert abs(b @ u - float(fx["bu"])) < 1e-7 * abs(float(fx["bu"]))
    assert inf["relres_true"] < 2e-8
    op.close()


def test_c2_solve_matches_oracle_fixture(gpu):
    """The headline workload at full size (192x96x96 H8 cantilever, 5.45M
    DOF, x = 0 face clamped, unit line load, Gaussian random field density
    seed 0; FP64 V(1,1) over 5 levels, relres 1e-8) against the oracle's FP64
    GMG-PCG on the same problem, committed by
    tests/golden/make_c2_oracle_fixture.py (about 75 s and 47 GiB of host
    memory): the same iteration count (+-1), the compliance f.u within 1e-8
    relative (the bar for C2), ||u|| within 1e-8 and u at 4096 seeded free
    DOFs within 1e-6."""
    import hashlib
    import torch
    from tests.conftest import GOLDEN
    from topoptlab_amd.fem import StiffnessOperator
    fx = np.load(GOLDEN / "c2_oracle_192x96x96.npz")
    nelx, nely, nelz = (int(v) for v in fx["dims"])
    say("c2 oracle fixture", (nelx, nely, nelz))
    ndof = 3 * (nelx + 1) * (nely + 1) * (nelz + 1)
    x = workloads.gaussian_random_density(nelx, nely, nelz, seed=0)
    assert hashlib.sha256(x.tobytes()).hexdigest() == str(fx["density_sha256"])
    f, fixed, free = fem.clamped_lineload_3d(nelx, nely, nelz, ndof)
    op = StiffnessOperator(nelx, nely, nelz, KE=fem.ke_elast_3d(), fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=int(fx["nlevels"]), precision="fp64")
    ft = torch.as_tensor(f[:, 0], device=gpu)
    u, inf = op.solve(ft, rtol=1e-8)
    c = float(torch.dot(ft, u).item())
    uf = u[op._free_t].cpu().numpy()
    c_ref = float(fx["compliance"])
    say("iterations", inf["iterations"], "oracle", int(fx["iterations"]),
        "compliance rel diff", abs(c - c_ref) / c_ref)
    assert abs(inf["iterations"] - int(fx["iterations"])) <= 1
    assert abs(c - c_ref) <= 1e-8 * abs(c_ref)
    assert abs(np.linalg.norm(uf) - float(fx["unorm"])) <= 1e-8 * float(fx["unorm"])
    assert rel(uf[fx["idx_free"]], fx["u_idx"]) < 1e-6
    assert inf["relres_true"] < 1e-8
    op.close()


@pytest.mark.parametrize("dims,ndpn,nlevels", [((32, 16, 16), 3, 4), ((64, 32, None), 2, 4)])
@pytest.mark.parametrize("gamma_from", [1, 2])
def test_pcg_wcycle_fp64_matches_oracle(gpu, gamma_from, dims, ndpn, nlevels):
    """FP64 W-cycle (gamma = 2 from level gamma_from): the first PCG iterate
    (alpha times the W-cycle applied to b) equals the oracle's, and PCG takes
    the same iteration count as the oracle's W-cycle preconditioner, with
    the same solution."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("gmg fp64 W-cycle", dims, nlevels, gamma_from)
    KE, edof, ndof, fixed, free, _ = _problem(dims, ndpn)
    x = workloads.gaussian_random_density(*dims[:2], dims[2], seed=5, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free).tocsr()
    b = np.random.default_rng(3).standard_normal(len(free))
    h = mg.Hierarchy(K, *dims, ndpn, fixed, nlevels)
    wcyc = lambda r: h.vcycle(r, gamma=2, gamma_from=gamma_from)
    u_ref, info, it_ref = mg.pcg(K, b, wcyc, rtol=1e-10)
    _, _, it_v = mg.pcg(K, b, h.vcycle, rtol=1e-10)
    assert info == 0 and it_ref <= it_v
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=nlevels, precision="fp64", gamma=2, gamma_from=gamma_from)
    x1_ref, _, _ = mg.pcg(K, b, wcyc, rtol=1e-10, maxiter=1)
    x1, _ = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10, maxiter=1,
                     raise_on_fail=False)
    assert rel(x1[op._free_t].cpu().numpy(), x1_ref) < 1e-8
    u, inf = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10)
    say("W-cycle iterations", inf["iterations"], "oracle", it_ref, "V", it_v)
    assert abs(inf["iterations"] - it_ref) <= 1, (inf, it_ref)
    assert rel(u[op._free_t].cpu().numpy(), u_ref) < 1e-7
    op.close()


# radii below one element (the element itself only) and wider than the grid;
# a one-element grid and a thin 3D grid
@pytest.mark.parametrize("dims", [(60, 20, None), (13, 7, None), (10, 8, 6), (1, 1, None),
                                  (4, 3, 2)])
@pytest.mark.parametrize("rmin", [2.4, 1.2, 3.0, 0.7, 5.5])
def test_filter_bitexact(gpu, dims, rmin):
    import torch
    from topoptlab_amd.filter import assemble_matrix_filter
    say("filter", dims, rmin)
    nelx, nely, nelz = dims
    H, Hs = fem.assemble_matrix_filter(nelx, nely, rmin, nelz)
    Hd, Hsd = assemble_matrix_filter(nelx, nely, rmin, nelz, device=gpu)
    assert np.array_equal(Hsd.cpu().numpy(), Hs)
    n = H.shape[0]
    rng = np.random.default_rng(0)
    x = rng.random(n)
    dc = -rng.random(n)
    xt = torch.as_tensor(x, device=gpu)
    dct = torch.as_tensor(dc, device=gpu)
    assert np.array_equal(Hd.apply(xt, 0).cpu().numpy(), H @ x / Hs[:, 0])
    assert np.array_equal(Hd.apply(dct, 1).cpu().numpy(), H @ (dc / Hs[:, 0]))
    assert np.array_equal((Hd @ xt).cpu().numpy(), H @ x)
    ref2 = np.asarray((H @ (x * dc)) / Hs[:, 0]) / np.maximum(0.001, x)
    assert np.array_equal(Hd.apply(dct, 2, x=xt).cpu().numpy(), ref2)


@pytest.mark.parametrize("dims,ndpn,ncases", [((12, 8, None), 2, 2), ((6, 4, 5), 3, 1),
                                              ((10, 6, None), 1, 1), ((1, 1, None), 2, 2),
                                              ((1, 1, 1), 3, 1), ((3, 1, 2), 1, 3)])
def test_compliance_sens(gpu, dims, ndpn, ncases):
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("compliance", dims, ndpn, ncases)
    KE, edof, ndof, fixed, free, x = _problem(dims, ndpn)
    u = np.random.default_rng(4).standard_normal((ndof, ncases))
    obj_ref, dc_ref = fem.compliance_sens(KE, edof, u, x)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    obj, dc = op.compliance_sens(torch.as_tensor(u.T.copy(), device=gpu), torch.as_tensor(x, device=gpu))
    assert obj == pytest.approx(obj_ref, rel=1e-13)
    assert rel(dc.cpu().numpy(), dc_ref) < 1e-13
    op.close()


# one element, and more elements than the bisection grid has threads (1024 x 256)
@pytest.mark.parametrize("n,g", [(1200, 0.0), (5000, 0.3), (70000, -0.01), (1, 0.0),
                                 (300000, 0.05)])
def test_oc_bitexact(gpu, n, g):
    import torch
    from topoptlab_amd.optimizer import oc_top88
    say("oc", n, g)
    rng = np.random.default_rng(n)
    x = rng.random(n)
    dc = -rng.random(n) * 10 ** rng.uniform(-3, 1, n)
    dv = 0.5 + rng.random(n)
    xr, gr = fem.oc_top88(x, 0.5, dc, dv, g)
    xd, gd = oc_top88(torch.as_tensor(x, device=gpu), 0.5, torch.as_tensor(dc, device=gpu),
                      torch.as_tensor(dv, device=gpu), g)
    assert np.array_equal(xd.cpu().numpy(), xr)
    assert gd == pytest.approx(gr, rel=1e-12, abs=1e-12)


def test_main_mbb_trajectory(gpu):
    """Per-iteration objective of the device OC loop against the oracle's
    scipy-direct loop: C1 (60x20 MBB, ft=1), 100 iterations, no early stop."""
    from topoptlab_amd.topology_optimization import main
    say("main mbb trajectory")
    h_ref, h = [], []
    topopt.topopt_loop(60, 20, 0.5, 2.4, ft=1, nouteriter=100, stop_on_change=False, history=h_ref)
    main(60, 20, 0.5, 2.4, ft=1, nouteriter=100, stop_on_change=False, history=h,
         lin_solver_kw={"name": "amd-gmgpcg", "rtol": 1e-12}, device=gpu)
    objs = np.array([o for o, _, _ in h])
    objs_ref = np.array([o for o, _, _ in h_ref])
    err = np.abs(objs - objs_ref) / objs_ref
    print("max rel obj error over 100 iterations:", err.max())
    assert err.max() < 1e-8


@pytest.mark.parametrize("ft,ref_idx", [(1, 1), (0, 0)])
def test_main_mbb_golden(gpu, kats, ft, ref_idx):
    """tests/test_compliance.py:8-35 on the device path (5 decimals)."""
    from topoptlab_amd.topology_optimization import main
    say("main mbb golden", ft)
    ref = kats["compliance_filters"]["rows"][ref_idx][-1]
    x, obj = main(60, 20, 0.5, 2.4, ft=ft, lin_solver_kw={"name": "amd-gmgpcg", "rtol": 1e-12},
                  device=gpu)
    assert x.shape == (1200, 1)
    assert obj == pytest.approx(ref, abs=1.5e-5)


@pytest.mark.parametrize("ft,ref_idx", [(0, 2), (1, 3)])
def test_main_helmholtz_golden(gpu, kats, ft, ref_idx):
    """tests/test_compliance.py:11-12 (Helmholtz filter) on the device path."""
    from topoptlab_amd.topology_optimization import main
    say("main helmholtz golden", ft)
    ref = kats["compliance_filters"]["rows"][ref_idx][-1]
    x, obj = main(60, 20, 0.5, 2.4, ft=ft, filter_mode="helmholtz",
                  lin_solver_kw={"name": "amd-gmgpcg", "rtol": 1e-12}, device=gpu)
    assert x.shape == (1200, 1)
    assert obj == pytest.approx(ref, abs=1.5e-5)


# 2 GMG levels, Jacobi (odd sizes), one element, a strip two elements wide,
# and 4 GMG levels
@pytest.mark.parametrize("nelx,nely", [(60, 20), (37, 13), (1, 1), (2, 64), (256, 128)])
def test_helmholtz_filter_matches_oracle(gpu, nelx, nely):
    import torch
    from scipy.sparse.linalg import spsolve
    from topoptlab_amd.filter import HelmholtzFilter
    say("helmholtz filter", nelx, nely)
    KF, TF = fem.assemble_helmholtz_filter(nelx, nely, 2.4)
    x = np.random.default_rng(nelx).random(nelx * nely)
    ref = TF.T @ spsolve(KF, TF @ x)
    H = HelmholtzFilter(nelx, nely, 2.4, device=gpu)
    y = H.apply_filter(torch.as_tensor(x, device=gpu)).cpu().numpy()
    assert rel(y, ref) < 1e-10
    assert np.abs(H.apply_filter(torch.full((nelx * nely,), 0.3, dtype=torch.float64,
                                            device=gpu)).cpu().numpy() - 0.3).max() < 1e-10
    H.close()


def test_main_twoloads_golden(gpu, kats):
    from topoptlab_amd.topology_optimization import main
    say("main twoloads golden")
    ref = kats["compliance_filters"]["rows"][4][-1]
    _, obj = main(60, 20, 0.5, 2.4, ft=0, bcs=example_bc.cantilever_2d_twoloads,
                  lin_solver_kw={"name": "amd-gmgpcg", "rtol": 1e-12}, device=gpu)
    assert obj == pytest.approx(ref, abs=1.5e-5)


def test_c2_size_operator_properties(gpu):
    """192x96x96 H8 (5,447,811 DOF): symmetry, and rigid translations in the
    null space of the unconstrained operator."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("c2 properties")
    dims = (192, 96, 96)
    x = workloads.gaussian_random_density(*dims, seed=0)
    op = StiffnessOperator(*dims, KE=fem.ke_elast_3d(), fixed=[], device=gpu)
    op.set_density(x)
    g = torch.Generator(device=gpu).manual_seed(0)
    u = torch.randn(op.ndof, dtype=torch.float64, device=gpu, generator=g)
    v = torch.randn(op.ndof, dtype=torch.float64, device=gpu, generator=g)
    a = torch.dot(v, op.matvec_full(u)).item()
    b = torch.dot(u, op.matvec_full(v)).item()
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))
    for d in range(3):
        t = torch.zeros(op.ndof, dtype=torch.float64, device=gpu)
        t[d::3] = 1.0
        y = op.matvec_full(t)
        assert y.abs().max().item() < 1e-13
    # linearity of the fp32 path against the fp64 path
    y64 = op.matvec_full(u)
    y32 = op.matvec_full(u, "fp32")
    assert (torch.linalg.norm(y32 - y64) / torch.linalg.norm(y64)).item() < 1e-5
    # matrix-core operator at full size: error per input type
    for dtype, tol in (("bf16", 4e-3), ("fp16", 5e-4)):
        y16 = op.matvec_full(u, dtype)
        err = (torch.linalg.norm(y16 - y64) / torch.linalg.norm(y64)).item()
        say("c2 mfma rel err", dtype, err)
        assert err < tol
    op.close()


# MFMA (bf16/fp16 inputs, fp32 accumulation) fine-level operator of the
# V-cycle.  Error bound: each element product is (KE/kmax) (u_e/max|u_e|)
# rounded to the input type (unit roundoff 2^-9 bf16, 2^-12 fp16) and summed
# in fp32, so |y - Ku| <= c * u_in * sum_e s_e |KE| |u_e| per DOF; the tests
# check the 2-norm relative error at 3x the error observed on MI355X
# (bf16 1.0-1.4e-3, fp16 1.2-1.6e-4 over the cases below).
MFMA_TOL = {"bf16": 4e-3, "fp16": 5e-4}


@pytest.mark.parametrize("dims", [(6, 4, 5), (24, 16, 20), (20, 60, 7), (17, 9, 3)])
@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_mfma_ku_error(gpu, dims, dtype):
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    say("mfma K.u", dims, dtype)
    KE, edof, ndof, fixed, free, x = _problem(dims, 3, seed=4)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    u = np.random.default_rng(5).standard_normal(len(free))
    y_ref = K @ u
    yf = op.matvec_full(op.full(torch.as_tensor(u, device=gpu)), dtype).cpu().numpy()
    assert np.all(yf[fixed] == 0)
    err = rel(yf[free], y_ref)
    say("mfma K.u rel err", dtype, err)
    assert err < MFMA_TOL[dtype]
    # linear in u up to rounding: scaling u by a power of two scales y exactly
    y2 = op.matvec_full(op.full(torch.as_tensor(4.0 * u, device=gpu)), dtype).cpu().numpy()
    assert np.array_equal(y2, 4.0 * yf)
    # a unit-material KE scaled by 1e6 (stiff units) gives 1e6 x the result
    op2 = StiffnessOperator(*dims, KE=KE * 1048576.0, fixed=fixed, device=gpu)
    op2.set_density(x)
    y3 = op2.matvec_full(op.full(torch.as_tensor(u, device=gpu)), dtype).cpu().numpy()
    assert rel(y3, 1048576.0 * yf) < 1e-6
    op.close()
    op2.close()


@pytest.mark.parametrize("dtype", ["bf16", "fp16"])
def test_mfma_vcycle_pcg_converges(gpu, dtype):
    """The bf16/fp16 V-cycle preconditions PCG to the same FP64 tolerance."""
    import torch
    from topoptlab_amd.fem import StiffnessOperator
    dims = (16, 8, 8)
    say("mfma pcg", dims, dtype)
    KE, edof, ndof, fixed, free, _ = _problem(dims, 3)
    x = workloads.gaussian_random_density(*dims, seed=5, sigma=2.0)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free).tocsc()
    b = np.random.default_rng(3).standard_normal(len(free))
    ud = spsolve(K, b)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    op.setup_mg(nlevels=3, precision="fp32")
    _, inf32 = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10)
    op.setup_mg(nlevels=3, precision=dtype)
    u, inf = op.solve(op.full(torch.as_tensor(b, device=gpu)), rtol=1e-10)
    say("mfma pcg iterations", dtype, inf["iterations"], "fp32", inf32["iterations"])
    assert inf["relres_true"] < 2e-10
    assert inf["iterations"] <= 1.5 * inf32["iterations"] + 3
    assert rel(u[op._free_t].cpu().numpy(), ud) < 1e-6
    op.close()


@pytest.mark.parametrize("dims", [(6, 4, 5), (24, 16, 20), (17, 30, 9)])
def test_fp64_mfma_operator_matches_vector(gpu, dims):
    """The FP64 column walk on v_mfma_f64_16x16x4f64 (selectable for 3D H8
    elasticity) against the vector-FMA kernel (the default) and the assembled oracle: K.u
    to 1e-13, and GMG-PCG with either operator in every FP64 mode (fused
    product, V-cycle residual and Jacobi sweeps) to the same iteration count."""
    import torch
    from topoptlab_amd._lib import check, lib
    from topoptlab_amd.fem import StiffnessOperator
    say("fp64 mfma", dims)
    KE, edof, ndof, fixed, free, x = _problem(dims, 3)
    K = fem.apply_bc(fem.assemble_K(KE, edof, fem.simp(x), ndof), free)
    op = StiffnessOperator(*dims, KE=KE, fixed=fixed, device=gpu)
    op.set_density(x)
    u = torch.as_tensor(np.random.default_rng(4).standard_normal(len(free)), device=gpu)
    ys = {}
    for kind in (0, 1):
        check(lib().tl_set_fp64_operator(op._h, kind))
        ys[kind] = (op @ u).cpu().numpy()
        assert rel(ys[kind], K @ u.cpu().numpy()) < 1e-13, kind
    assert rel(ys[1], ys[0]) < 1e-14
    if all(d % 4 == 0 for d in dims):
        f = op.full(torch.as_tensor(np.random.default_rng(5).standard_normal(len(free)), device=gpu))
        res = {}
        for kind in (0, 1):
            check(lib().tl_set_fp64_operator(op._h, kind))
            op.setup_mg(nlevels=3, precision="fp64")
            s, inf = op.solve(f, rtol=1e-10)
            res[kind] = (s.cpu().numpy(), inf["iterations"])
            assert inf["relres_true"] < 2e-10
        assert abs(res[1][1] - res[0][1]) <= 1
        assert rel(res[1][0], res[0][0]) < 1e-7
    op.close()


# both asymptote branches; one design variable, and more variables than the
# reduction grid has threads (1024 x 256)
@pytest.mark.parametrize("it,n", [(0, 3000), (5, 3000), (0, 1), (5, 1), (5, 300000)])
def test_mma_update_matches_oracle(gpu, it, n):
    """tl_mma_update against the oracle's mmasub: both asymptote branches."""
    import torch
    from oracle import mma
    from topoptlab_amd.optimizer import mma_defaultkws, mma_update
    say("mma update", it, n)
    rng = np.random.default_rng(it)
    x = 0.05 + 0.9 * rng.random(n)
    xold1 = np.clip(x + 0.05 * rng.standard_normal(n), 0.0, 1.0)
    xold2 = np.clip(xold1 + 0.05 * rng.standard_normal(n), 0.0, 1.0)
    df0 = -rng.random(n) * 10.0
    dfdx = np.full(n, 1.0 / n)
    fval = 0.013
    low0, upp0 = x - 0.2 - 0.1 * rng.random(n), x + 0.2 + 0.1 * rng.random(n)
    okw = mma.mma_defaultkws(n, 1)
    out = mma.mmasub(1, n, it, x[:, None], okw["xmin"], okw["xmax"], xold1[:, None],
                     xold2[:, None], 0.0, df0[:, None], np.array([[fval]]), dfdx[None, :],
                     low0[:, None], upp0[:, None], okw["a0"], okw["a"], okw["c"], okw["d"],
                     move=okw["move"])
    kw = mma_defaultkws(n, 1, device=gpu)
    kw["low"] = torch.as_tensor(low0, device=gpu)
    kw["upp"] = torch.as_tensor(upp0, device=gpu)
    t = lambda a: torch.as_tensor(a, device=gpu)
    xd = mma_update(t(x), t(xold1), t(xold2), t(df0), fval, t(dfdx), kw, iter=it)
    assert rel(kw["low"].cpu().numpy(), out[9][:, 0]) < 1e-15
    assert rel(kw["upp"].cpu().numpy(), out[10][:, 0]) < 1e-15
    assert rel(xd.cpu().numpy(), out[0][:, 0]) < 1e-9
    assert kw["newton_its"] > 0


def test_main_mma_golden(gpu, kats):
    """tests/test_compliance.py:62-86 (MMA) on the device path, 5 decimals."""
    from topoptlab_amd.topology_optimization import main
    say("main mma golden")
    (opt, ref), = kats["compliance_optimizers"]["rows"]
    x, obj = main(60, 20, 0.5, 2.4, ft=1, optimizer="mma",
                  lin_solver_kw={"name": "amd-gmgpcg", "rtol": 1e-12}, device=gpu)
    assert x.shape == (1200, 1)
    assert obj == pytest.approx(ref, abs=1.5e-5)
