"""mdg_stokes_residual_static, the entry point that allocates nothing: its
residual F and norms equal mdg_stokes_residual's bit for bit, on two calls
in a row that share the device's static scratch, the second on a smaller
grid; host arrays are refused."""
import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S

pytestmark = pytest.mark.gpu


def _case(O, n, seed):
    box = (-0.5, 0.5, -0.5, 0.5)
    g = O.Grid(n, n, *box)
    g.pure_shear(1.0)
    ctx = mdgpu.Context(n, n, *box)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    rng = np.random.default_rng(seed)
    nc, nv = (n - 1) ** 2, n * n
    host = dict(sxxd=rng.standard_normal(nc), szzd=rng.standard_normal(nc), sxz=rng.standard_normal(nv),
                p_corr=rng.standard_normal(nc), div=1e-3 * rng.standard_normal(nc),
                roger_x=rng.standard_normal(n * (n + 1)), roger_z=rng.standard_normal(n * (n + 1)))
    neq = ctx.eqn_count()[0]
    d = 1.0 + rng.random(neq)
    return ctx, host, d, neq


@pytest.mark.parametrize("scaled", [False, True])
def test_static_residual_matches_the_allocating_one(oracle_mod, scaled):
    for n, seed in ((257, 1), (65, 2)):          # second call: a smaller grid, same scratch
        ctx, host, d, neq = _case(oracle_mod, n, seed)
        dev = {k: ctx.device(v.size, host=v) for k, v in host.items()}
        dd = ctx.device(neq, host=d) if scaled else None
        F_ref, n_ref = ctx.stokes_residual(*[dev[k] for k in ("sxxd", "szzd", "sxz", "p_corr", "div", "roger_x",
                                                              "roger_z")], d=dd, F=ctx.device(neq))
        F_ref = F_ref.numpy()
        Fs = ctx.device(neq, host=np.full(neq, np.nan))
        n_st = ctx.stokes_residual_static(*[dev[k] for k in ("sxxd", "szzd", "sxz", "p_corr", "div", "roger_x",
                                                             "roger_z")], Fs, d=dd)
        assert np.array_equal(Fs.numpy(), F_ref), n
        assert np.array_equal(n_st, np.asarray(n_ref)), (n, n_st, n_ref)
        with pytest.raises(TypeError):
            ctx.stokes_residual_static(host["sxxd"], *[dev[k] for k in ("szzd", "sxz", "p_corr", "div",
                                                                       "roger_x", "roger_z")], Fs)
        ctx.close()
