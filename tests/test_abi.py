"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes

import mdgpu


def test_library_exports_every_declared_function():
    names = mdgpu.declared_functions()
    assert "mdg_p2mastah" in names and "mdg_create" in names
    L = ctypes.CDLL(mdgpu.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    L = mdgpu.lib()
    for n in mdgpu.declared_functions():
        assert getattr(L, n).argtypes is not None, n
