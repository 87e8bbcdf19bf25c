"""The SIMP/OC optimisation loop on the GPU with topoptlab's ``main`` signature
(topoptlab/topology_optimization.py:47-688).

Supported branch (the hot path): compliance objective, modified SIMP,
``optimizer="oc"`` (oc_top88), ``ft`` in {-1, 0, 1} with ``filter_mode`` in
{"matrix", "convolution"} (both are the same distance filter), structured
Q4/H8 grids with unit elements, no body forces or springs; element flags
(passive 1 / active 2) with the OC optimizer.  Other options raise
``NotImplementedError``.

Per iteration everything stays on the device: SIMP scale, matrix-free
GMG-PCG solve per load case (warm started from the previous displacement,
``rhs0 = u[free]`` at topology_optimization.py:455), fused compliance +
sensitivity kernel, filter stencils and the device OC bisection; only the
scalars of the log line cross to the host.  With ``output_kw["write_log"]``
the header of topology_optimization.py:180-199 and one line per iteration
(669-670) are written by ``SimpleLogger`` (log_utils.py:147-350), so
``parse_simple_logfile`` reads the file back.
"""
from __future__ import annotations

from typing import Callable, Dict, List, Optional

import numpy as np
import torch

from .elements import lk_linear_elast_2d, lk_linear_elast_3d
from .example_bc import mbb_2d
from .fem import StiffnessOperator
from .filter import DeviceFilter, HelmholtzFilter
from .log_utils import EmptyLogger, SimpleLogger
from .optimizer import mean_maxdiff, mma_defaultkws, mma_update, oc_top88

LOG_LINE = "it.: {0} obj.: {1:.10f} vol.: {2:.10f} ch.: {3:.10f}"
FILTER_NAMES = ["Sensitivity based", "Density based", "Haeviside Guest",
                "Haeviside complement Sigmund 2007", "Haeviside eta projection",
                "Volume Preserving eta projection", "No filter"]


def default_outputkw() -> Dict:
    return {"file": "topopt", "display": False, "export": False, "write_log": False,
            "profile": False, "debug": 0, "verbosity": 20}


def main(nelx: int, nely: int, volfrac: float, rmin: float, ft: int = 1,
         filter_kw: Dict = {}, simulation_kw: Dict = {"grid": "regular", "element order": 1,
                                                      "meshfile": None},
         nelz: Optional[int] = None, filter_mode: str = "matrix",
         lin_solver_kw: Dict = {"name": "amd-gmgpcg"},
         preconditioner_kw: Dict = {"name": None}, assembly_mode: str = "full",
         materials_kw: Dict = {"E": 1.0}, body_forces_kw: Dict = {},
         bcs: Callable = mbb_2d, lk: Optional[Callable] = None, l=1.0,
         obj_func=None, obj_kw: Dict = {}, matinterpol=None, matinterpol_dx=None,
         matinterpol_kw: Dict = {"eps": 1e-9, "penal": 3.0},
         el_flags=None, optimizer: str = "oc", optimizer_kw: Optional[Dict] = None,
         mix=None, accelerator_kw: Optional[Dict] = None, nouteriter: int = 2000,
         ninneriter: int = 15, output_kw: Dict = default_outputkw(),
         stop_on_change: bool = True, history: Optional[List] = None, device=None):
    """Run the OC loop; return ``(x, obj)`` like topology_optimization.py:688
    (x as a float64 numpy array of shape (nel, 1))."""
    if optimizer not in ("oc", "mma"):
        raise NotImplementedError(f"optimizer {optimizer!r} is not on the device path")
    if ft not in (-1, 0, 1):
        raise NotImplementedError(f"filter {ft!r} is not on the device path")
    if filter_mode not in ("matrix", "convolution", "helmholtz"):
        raise NotImplementedError(f"filter_mode {filter_mode!r} is not on the device path")
    helm = filter_mode == "helmholtz" and ft in (0, 1)
    if body_forces_kw or mix is not None:
        raise NotImplementedError("body forces and mixing are not supported")
    if el_flags is not None and optimizer != "oc":
        # topology_optimization.py:248-259 gives passive elements x = 1 with
        # bounds [0, 1e-9] under MMA, a subproblem with empty boxes
        raise NotImplementedError("element flags are supported with optimizer='oc'")
    if (accelerator_kw or {}).get("accelerator") is not None:
        raise NotImplementedError("accelerators are not supported")
    if simulation_kw.get("grid", "regular") != "regular":
        raise ValueError("only regular grids are supported")
    solver = dict(lin_solver_kw)
    name = solver.pop("name", "amd-gmgpcg")
    if name not in ("amd-gmgpcg", "amd-pcg"):
        raise ValueError("Unknown solver: ", name)
    solver.setdefault("rtol", 1e-10)
    solver.setdefault("maxiter", 20000)
    mg_kw = {k: v for k, v in preconditioner_kw.items() if k != "name"}
    mg_kw.setdefault("precision", "fp64")
    eps = float(matinterpol_kw.get("eps", 1e-9))
    penal = float(matinterpol_kw.get("penal", 3.0))
    ndim = 2 if nelz is None else 3
    if lk is None:
        lk = lk_linear_elast_2d if ndim == 2 else lk_linear_elast_3d
    KE = np.asarray(lk(l=np.full(ndim, float(l)) if np.isscalar(l) else l), dtype=np.float64)
    ndpn = KE.shape[-1] // 2 ** ndim
    nn = int(np.prod(np.array([nelx, nely, nelz][:ndim]) + 1))
    ndof = ndpn * nn
    n = int(np.prod([nelx, nely, nelz][:ndim]))
    if ndim == 2:
        u_h, f_h, fixed, free, springs = bcs(nelx=nelx, nely=nely, ndof=ndof)
    else:
        u_h, f_h, fixed, free, springs = bcs(nelx=nelx, nely=nely, nelz=nelz, ndof=ndof)
    if springs:
        raise NotImplementedError("springs are not supported")
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    out_kw = dict(default_outputkw(), **output_kw)
    if out_kw["write_log"]:  # topology_optimization.py:180-199
        log = SimpleLogger(file=out_kw["file"], verbosity=out_kw["verbosity"])
        log.info(f"optimizer {optimizer}")
        log.info(f"number of spatial dimensions: {ndim}")
        log.info("elements: " + " x ".join([f"{nelx}", f"{nely}", f"{nelz}"][:ndim]))
        log.info(f"volfrac: {volfrac} rmin: {rmin}" if volfrac is not None else f"rmin: {rmin}")
        log.info("filter: " + FILTER_NAMES[ft])
        log.info(f"filter mode: {filter_mode}")
    else:
        log = EmptyLogger()
    op = StiffnessOperator(nelx, nely, nelz, KE, fixed, eps, penal, device=dev)
    if helm:  # topology_optimization.py:360-365
        H = HelmholtzFilter(nelx, nely, rmin, nelz, device=dev)
    else:
        H = DeviceFilter(nelx, nely, rmin, nelz, device=dev) if ft in (0, 1) else None
    f = torch.as_tensor(np.ascontiguousarray(f_h.T), dtype=torch.float64, device=dev)  # [nbc][ndof]
    ncases = f.shape[0]
    u = torch.zeros_like(f)
    x = torch.full((n,), float(volfrac), dtype=torch.float64, device=dev)
    xPhys = x.clone()
    g = 0.0
    flags = None
    if el_flags is not None:
        flags = torch.as_tensor(np.asarray(el_flags).reshape(-1), dtype=torch.int32, device=dev)
        if flags.numel() != n:
            raise ValueError("el_flags must have one entry per element")
    mma = optimizer == "mma"
    if mma:  # topology_optimization.py:244-246; mma_utils.py:8-60
        mkw = mma_defaultkws(n, 1, ft, device=dev)
        mkw.update({k: v for k, v in (optimizer_kw or {}).items() if k not in ("nconstr",)})
        for k in ("xmin", "xmax", "low", "upp"):
            mkw[k] = torch.as_tensor(np.asarray(mkw[k].cpu() if torch.is_tensor(mkw[k]) else mkw[k],
                                                dtype=np.float64).reshape(-1), device=dev).clone()
        xhist = [x.clone() for _ in range(3)]
    # constraint gradient (topology_optimization.py:517-520): 1/n for MMA, 1 for OC
    dv0 = torch.full((n,), 1.0 / n if mma else 1.0, dtype=torch.float64, device=dev)
    if helm:
        dv_f = H.apply_filter(dv0) if ft == 1 else dv0
    else:
        dv_f = H.apply(dv0, 1) if ft == 1 else dv0
    obj = 0.0
    for loop in range(nouteriter):
        op.set_density(xPhys)
        if name == "amd-gmgpcg":
            op.setup_mg(**mg_kw)
        prec = "gmg" if name == "amd-gmgpcg" else "jacobi"
        for i in range(ncases):
            u[i], _ = op.solve(f[i], u0=u[i], preconditioner=prec, **solver)
        obj, dc = op.compliance_sens(u, xPhys)
        if helm and ft == 0:  # topology_optimization.py:554-555
            dc = H.apply_filter(dc * xPhys) / torch.clamp(x, min=0.001)
        elif helm:  # 568-570
            dc = H.apply_filter(dc)
        elif ft == 0:
            dc = H.apply(dc, 2, x=x)
        elif ft == 1:
            dc = H.apply(dc, 1)
        xold = x
        if mma:
            # mmasub(iter=i) at topology_optimization.py:592-602: i is the
            # load-case loop's last index (463), 0 for one load case, so the
            # asymptotes restart every iteration; kept for identical results
            fval = float(mean_maxdiff(xPhys, None)[0]) - volfrac
            x = mma_update(x, xhist[-1], xhist[-2], dc, fval, dv_f, mkw, iter=ncases - 1)
            xhist = (xhist + [x.clone()])[-4:]
        else:
            x, g = oc_top88(x, volfrac, dc, dv_f, g, el_flags=flags)
        if helm:  # 646-647
            xPhys = H.apply_filter(x) if ft == 1 else x.clone()
        else:
            xPhys = H.apply(x, 0) if ft == 1 else x.clone()
        vol, change = mean_maxdiff(xPhys, None)[0], mean_maxdiff(x, xold)[1]
        if history is not None:
            history.append((obj, vol, change))
        log.info(LOG_LINE.format(loop + 1, obj, vol, change))
        if stop_on_change and change < 0.01:
            break
    if out_kw["export"]:  # topology_optimization.py:675-679
        from .output_designs import export_vtk
        export_vtk(out_kw["file"], nelx, nely, xPhys.cpu().numpy(), nelz,
                   x=x.cpu().numpy(), u=u.T.cpu().numpy(), f=f.T.cpu().numpy(), volfrac=volfrac)
    op.close()
    if H is not None:
        H.close()
    return x.cpu().numpy()[:, None], obj
