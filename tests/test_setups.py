"""Host-side set-up helpers of the package against the oracle's boundary tags."""
import numpy as np

from mdgpu import learned


def test_free_slip_tags_match_oracle_pure_shear(oracle_mod):
    for Nx in (5, 17):
        g = oracle_mod.Grid(Nx, Nx, 0.0, 1.0, 0.0, 1.0)
        g.pure_shear(0.0)
        BCu_t, BCu_v, BCv_t, BCv_v, BCp_t, BCg_t = learned.free_slip_bc(Nx)
        assert np.array_equal(BCu_t, g.BCu_t) and np.array_equal(BCv_t, g.BCv_t)
        assert np.array_equal(BCp_t, g.BCp_t) and np.array_equal(BCg_t, g.BCg_t)
        assert not BCu_v.any() and not BCv_v.any()
        assert np.array_equal(g.BCu_v, BCu_v) and np.array_equal(g.BCv_v, BCv_v)


def test_init_params_layout():
    p = learned.init_params(0)
    assert p.dtype == np.float32
    assert p.size == sum(9 * ci * co + co for ci, co in learned.LAYERS)
