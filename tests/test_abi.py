"""The C-ABI library loads without a GPU and exports every declared entry point."""
import ctypes

import pytest

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


def test_rheology_inputs_are_checked_before_any_call():
    """The rheology wrappers hand raw pointers to the library: a node array
    of the wrong length is refused on the host."""
    import numpy as np
    import mdgpu
    n, nph = 12, 2
    f, keep = mdgpu.Context._rheo_inputs(dict(perc=np.ones((nph, n)), T=np.ones(n), d=np.ones(n, np.float32)), n, nph)
    assert f.perc and f.T and f.d and not f.phi and len(keep) == 3
    assert keep[2].dtype == np.float64            # converted, not reinterpreted
    with pytest.raises(ValueError, match="phi holds 11 values"):
        mdgpu.Context._rheo_inputs(dict(perc=np.ones((nph, n)), phi=np.ones(n - 1)), n, nph)
    with pytest.raises(ValueError, match="perc"):
        mdgpu.Context._rheo_inputs(dict(perc=np.ones(n)), n, nph)
