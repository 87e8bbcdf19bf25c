"""tl_oc_update_noalloc (statically allocated scratch) against tl_oc_update
through the C ABI: bit-identical designs and gt on two calls in a row, the
second with a smaller input and different element flags (marked ``gpu``)."""
import ctypes as C
import sys

import numpy as np
import pytest

from topoptlab_amd import _lib

pytestmark = pytest.mark.gpu


def _call(fn, x, dc, dv, flags, g, xnew, stream):
    gt = C.c_double()
    p = lambda t: C.c_void_p(0 if t is None else t.data_ptr())
    _lib.check(fn(x.numel(), p(x), p(dc), p(dv), p(flags), float(g), 0.2, 0.0, 1e9, p(xnew),
                  C.byref(gt), C.c_void_p(stream)))
    return gt.value


def test_noalloc_matches_allocating_entry_point(gpu):
    import torch
    lib = _lib.lib()
    stream = torch.cuda.current_stream(gpu).cuda_stream
    rng = np.random.default_rng(7)
    for n, with_flags in [(300_001, True), (4_097, False), (1_769_472, True), (1, False)]:
        print(f"[case] noalloc oc update n={n} flags={with_flags}", flush=True)
        sys.stdout.flush()
        t = lambda a: torch.as_tensor(a, device=gpu)
        x = t(rng.random(n))
        dc = t(-rng.random(n) * 10 ** rng.uniform(-3, 1, n))
        dv = t(0.5 + rng.random(n))
        flags = t(rng.integers(0, 3, n).astype(np.int32)) if with_flags else None
        a = torch.full((n,), -1.0, dtype=torch.float64, device=gpu)
        b = torch.full((n,), -2.0, dtype=torch.float64, device=gpu)
        ga = _call(lib.tl_oc_update, x, dc, dv, flags, 0.05, a, stream)
        gb = _call(lib.tl_oc_update_noalloc, x, dc, dv, flags, 0.05, b, stream)
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert ga == gb
        if with_flags:
            f = flags.cpu().numpy()
            bb = b.cpu().numpy()
            assert (bb[f == 1] == 0).all() and (bb[f == 2] == 1).all()


def test_noalloc_rejects_bad_arguments(gpu):
    lib = _lib.lib()
    assert lib.tl_oc_update_noalloc(0, None, None, None, None, 0.0, 0.2, 0.0, 1e9, None, None,
                                    None) == _lib.TL_EINVAL
