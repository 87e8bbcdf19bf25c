"""CPU checks of the C ABI: the library builds for gfx950, loads, and exports
every entry point that include/topoptlab_amd.h declares (no GPU calls)."""
import ctypes
import re

import pytest

from tests.conftest import ROOT
from topoptlab_amd import _lib


def declared():
    txt = (ROOT / "include" / "topoptlab_amd.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tl_[A-Za-z_0-9]+)\s*\(", txt)))


def test_header_declarations_are_bound():
    names = declared()
    assert len(names) >= 15
    assert set(names) == set(_lib.EXPORTED)


def test_library_exports_every_symbol():
    lib = _lib.lib()
    for name in declared():
        assert hasattr(lib, name), name


def test_error_paths_without_gpu():
    lib = _lib.lib()
    # argument validation happens before any device work
    rc = lib.tl_create(None, ctypes.byref(ctypes.c_void_p()))
    assert rc == _lib.TL_EINVAL
    assert b"null" in lib.tl_last_error()
    with pytest.raises(ValueError):
        _lib.check(rc, "tl_create")
    desc = _lib.GridDesc(4, 1, 1, 1, 1, None, 1e-9, 3.0, None, 0, None)
    assert lib.tl_create(ctypes.byref(desc), ctypes.byref(ctypes.c_void_p())) == _lib.TL_EINVAL
    assert b"ndim" in lib.tl_last_error()
    h = ctypes.c_void_p()
    assert lib.tl_filter_create(2, 4, 4, 0, -1.0, None, ctypes.byref(h)) == _lib.TL_EINVAL
    assert lib.tl_oc_update(0, None, None, None, None, 0.0, 0.2, 0.0, 1e9, None, None,
                            None) == _lib.TL_EINVAL
