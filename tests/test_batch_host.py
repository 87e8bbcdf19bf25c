"""CPU checks of the batched workload generator and of the batch entry
points' argument validation (no GPU calls)."""
import ctypes

import numpy as np

from topoptlab_amd import _lib
from topoptlab_amd.fem import auto_levels
from topoptlab_amd.workloads import batched_instances, hierarchy_halvings


def test_batched_instances_shapes():
    a = batched_instances(seed=0)
    b = batched_instances(seed=0)
    for x, y in zip(a, b):  # deterministic
        assert (x["nelx"], x["nely"], x["density_seed"]) == (y["nelx"], y["nely"], y["density_seed"])
        assert np.array_equal(x.get("load_nodes", []), y.get("load_nodes", []))
    assert [d["kind"] for d in a] == ["elast"] * 32 + ["heat"] * 8
    for d in a:
        dx, dy = d["drawn"]
        assert 128 <= dx <= 1024 and 64 <= dy <= 512
        # rounded up to the multiple its own hierarchy needs, never below the draw
        ndpn = 2 if d["kind"] == "elast" else 1
        m = 1 << hierarchy_halvings(dx, dy, ndpn, 3000)
        assert d["nelx"] % m == 0 and d["nely"] % m == 0
        assert dx <= d["nelx"] < dx + m and dy <= d["nely"] < dy + m
        levels = auto_levels(d["nelx"], d["nely"], None, ndpn, 3000)
        f = 1 << (levels - 1)
        assert levels >= 2 and ndpn * (d["nelx"] // f + 1) * (d["nely"] // f + 1) <= 3000
        if d["kind"] == "elast":
            nodes = d["load_nodes"]
            assert len(nodes) == 4 and len(set(nodes.tolist())) == 4
            ix, iy = nodes // (d["nely"] + 1), nodes % (d["nely"] + 1)
            assert np.all(ix > 0)  # never on the clamped edge
            assert np.all((ix == d["nelx"]) | (iy == 0) | (iy == d["nely"]))
    sizes = {(d["nelx"], d["nely"]) for d in a}
    assert len(sizes) > 20  # uneven


def test_batch_entry_points_validate_arguments():
    lib = _lib.lib()
    opts = _lib.SolveOpts(1e-8, 0.0, 100, 0, 2, 8)
    assert lib.tl_batch_solve(0, None, None, None, None, ctypes.byref(opts), None, None) == _lib.TL_EINVAL
    assert b"empty" in lib.tl_last_error()
    assert lib.tl_set_tile2d(None, 0, 16) == _lib.TL_EINVAL
    assert lib.tl_autotune(None, 5, None, None) == _lib.TL_EINVAL
