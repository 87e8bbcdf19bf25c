"""ctypes binding of libtopoptlab_amd.so (the C ABI of include/topoptlab_amd.h).

The library is loaded from this package directory.  There is no fallback:
if it is missing or fails to load, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

TL_OK, TL_EINVAL, TL_ENOCONV, TL_EHIP = 0, 1, 2, 3
TL_PREC_FP64, TL_PREC_FP32, TL_PREC_BF16, TL_PREC_FP16 = 0, 1, 2, 3
TL_PRECOND_NONE, TL_PRECOND_JACOBI, TL_PRECOND_GMG = 0, 1, 2

LIB_PATH = Path(__file__).resolve().parent / "libtopoptlab_amd.so"


class GridDesc(C.Structure):
    _fields_ = [("ndim", C.c_int32), ("nelx", C.c_int32), ("nely", C.c_int32),
                ("nelz", C.c_int32), ("ndpn", C.c_int32),
                ("KE", C.POINTER(C.c_double)), ("eps", C.c_double),
                ("penal", C.c_double), ("fixed", C.POINTER(C.c_int64)),
                ("nfixed", C.c_int64), ("stream", C.c_void_p)]


class MgOpts(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("omega", C.c_double), ("nu1", C.c_int32),
                ("nu2", C.c_int32), ("precision", C.c_int32), ("gamma", C.c_int32),
                ("gamma_from", C.c_int32)]


class SolveOpts(C.Structure):
    _fields_ = [("rtol", C.c_double), ("atol", C.c_double), ("maxiter", C.c_int32),
                ("x0_given", C.c_int32), ("precond", C.c_int32),
                ("check_every", C.c_int32)]


class BatchOpts(C.Structure):
    _fields_ = [("nstreams", C.c_int32), ("chunk", C.c_int32), ("use_graphs", C.c_int32),
                ("lockstep", C.c_int32)]


class MmaOpts(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("move", "asyinit", "asydecr", "asyincr", "asymin",
                                          "asymax", "raa0", "albefa", "a0", "a", "c", "d")]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("status", C.c_int32),
                ("relres_recursive", C.c_double), ("relres_true", C.c_double)]


# name -> (restype, argtypes)
_SIGS = {
    "tl_create": (C.c_int, [C.POINTER(GridDesc), C.POINTER(C.c_void_p)]),
    "tl_destroy": (None, [C.c_void_p]),
    "tl_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tl_last_error": (C.c_char_p, []),
    "tl_ndof": (C.c_int64, [C.c_void_p]),
    "tl_nel": (C.c_int64, [C.c_void_p]),
    "tl_set_density": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tl_apply_K": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "tl_diag_K": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tl_setup_mg": (C.c_int, [C.c_void_p, C.POINTER(MgOpts)]),
    "tl_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(SolveOpts),
                           C.POINTER(SolveInfo)]),
    "tl_batch_solve": (C.c_int, [C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.POINTER(SolveOpts), C.POINTER(BatchOpts),
                                 C.POINTER(SolveInfo)]),
    "tl_set_tile2d": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32]),
    "tl_set_fp64_operator": (C.c_int, [C.c_void_p, C.c_int32]),
    "tl_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p]),
    "tl_autotune_clear": (None, []),
    "tl_compliance_sens": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p,
                                     C.POINTER(C.c_double), C.c_void_p]),
    "tl_filter_create": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double,
                                   C.c_void_p, C.POINTER(C.c_void_p)]),
    "tl_filter_destroy": (None, [C.c_void_p]),
    "tl_filter_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tl_filter_hs": (C.c_int, [C.c_void_p, C.c_void_p]),
    "tl_filter_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "tl_helmholtz_transfer": (C.c_int, [C.c_int32, C.c_int32, C.c_void_p, C.c_void_p, C.c_int32,
                                        C.c_void_p]),
    "tl_oc_update": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                               C.c_double,
                               C.c_double, C.c_double, C.c_double, C.c_void_p,
                               C.POINTER(C.c_double), C.c_void_p]),
    "tl_oc_update_noalloc": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_double, C.c_double, C.c_double,
                                       C.c_double, C.c_void_p, C.POINTER(C.c_double),
                                       C.c_void_p]),
    "tl_profile": (C.c_int, [C.c_void_p, C.c_int32]),
    "tl_profile_read": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(C.c_double),
                                  C.POINTER(C.c_int64)]),
    "tl_nn_conv3d": (C.c_int, [C.c_int32] * 7 + [C.c_void_p] * 5),
    "tl_nn_conv3d_bwd_data": (C.c_int, [C.c_int32] * 6 + [C.c_void_p] * 3 + [C.c_int32,
                                                                              C.c_void_p]),
    "tl_nn_conv3d_bwd_weight": (C.c_int, [C.c_int32] * 6 + [C.c_void_p] * 5),
    "tl_nn_relu_mask": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    "tl_nn_upsample": (C.c_int, [C.c_int32] * 4 + [C.c_void_p, C.c_void_p, C.c_int32, C.c_void_p]),
    "tl_nn_upsample_t": (C.c_int, [C.c_int32] * 4 + [C.c_void_p, C.c_void_p, C.c_int32,
                                                     C.c_void_p]),
    "tl_nn_mse": (C.c_int, [C.c_int64] + [C.c_void_p] * 6),
    "tl_nn_adam": (C.c_int, [C.c_int64] + [C.c_void_p] * 5 + [C.c_int32, C.c_float, C.c_void_p]),
    "tl_nn_features": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "tl_nn_inject": (C.c_int, [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    "tl_nn_prolong": (C.c_int, [C.c_void_p] * 3),
    "tl_energy_scale": (C.c_int, [C.c_void_p] * 4 + [C.POINTER(C.c_double)]),
    "tl_mma_update": (C.c_int, [C.c_int64] + [C.c_void_p] * 6 + [C.c_double, C.c_void_p,
                                                                   C.c_void_p, C.c_void_p,
                                                                   C.c_int32, C.POINTER(MmaOpts),
                                                                   C.c_void_p, C.c_void_p]),
    "tl_mean_maxdiff": (C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.c_void_p]),
}

EXPORTED = tuple(_SIGS)

_lib = None


class LibraryMissing(RuntimeError):
    pass


def lib():
    """Load (once) and return the shared library; raise if it is unavailable."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise LibraryMissing(
                f"{LIB_PATH} is not built; run `python -m topoptlab_amd.build`")
        handle = C.CDLL(os.fspath(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str = "") -> None:
    """Translate a C-ABI status into the reference's Python exceptions."""
    if rc == TL_OK:
        return
    msg = lib().tl_last_error().decode(errors="replace")
    text = f"{what}: {msg}" if what else msg
    if rc == TL_EINVAL:
        raise ValueError(text)
    raise RuntimeError(text)
