"""GPU subgrid-diffusion marker updates (mdg_update_particle_stress_subgrid,
mdg_update_particle_pressure_subgrid) against the oracle's restatement of
UpdateParticleStress (subgrid_diffusion = 2, RheologyParticles.c:1811-1895)
and UpdateParticlePressure (:1621-1695).  The grid->marker steps are
bit-exact; the marker->grid step (P2Mastah) sums in another order, so the
results agree to 1e-12 of each field's scale.  Cases: both interpolation
stencils, inactive markers, device-resident markers, a 255^2 box, and the
no-relaxation limit, which must equal the plain (non-subgrid) GPU update."""
import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S

pytestmark = pytest.mark.gpu


def _case(O, n, seed):
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    rng = np.random.default_rng(seed)
    ph = rng.integers(0, 3, x.size).astype(np.int32)
    ph[::41] = -1
    nc, nv = n * n, (n + 1) ** 2
    f = {k: rng.standard_normal(nc) for k in ("sxxd", "szzd", "sxxd0", "szzd0", "p_in", "p0_n")}
    f.update({k: rng.standard_normal(nv) for k in ("sxz", "sxz0")})
    f["eta_s"] = rng.uniform(0.5, 2.0, nv)
    f["eta_n"] = rng.uniform(0.5, 2.0, nc)
    m = {k: rng.standard_normal(x.size) for k in ("sxx", "szz", "sxz", "P")}
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    ctx.markers_set(x, z, ph)
    return g, ctx, x, z, ph, f, m


def _close(a, b):
    scale = max(np.abs(b).max(), 1.0)
    assert np.abs(a - b).max() <= 1e-12 * scale, np.abs(a - b).max() / scale


@pytest.mark.parametrize("n,stencil,device", [(24, 1, False), (24, 9, False), (31, 1, True), (255, 1, False)])
def test_subgrid_matches_oracle(oracle_mod, n, stencil, device):
    O = oracle_mod
    g, ctx, x, z, ph, f, m = _case(O, n, seed=n + stencil)
    G, beta, dt = [1.5, 0.7, 3.0], [0.4, 2.0, 1.1], 0.9
    exp_s = O.update_particle_stress_subgrid(g, dt, G, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"],
                                             f["sxz0"], f["eta_s"], x, z, ph, m["sxx"], m["szz"], m["sxz"],
                                             stencil=stencil)
    exp_p, exp_dp = O.update_particle_pressure_subgrid(g, dt, beta, f["p_in"], f["p0_n"], f["eta_n"], x, z, ph, m["P"])
    out = [m[k].copy() for k in ("sxx", "szz", "sxz")]
    P = m["P"].copy()
    if device:
        dev = [mdgpu.DeviceArray(ctx, a.size, host=a) for a in out]
        dP = mdgpu.DeviceArray(ctx, P.size, host=P)
        ctx.update_particle_stress_subgrid(dt, G, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"], f["sxz0"],
                                           f["eta_s"], x, z, ph, *dev, interp_stencil=stencil)
        dp = ctx.update_particle_pressure_subgrid(dt, beta, f["p_in"], f["p0_n"], f["eta_n"], x, z, ph, dP)
        out = [d.numpy() for d in dev]
        P = dP.numpy()
    else:
        ctx.update_particle_stress_subgrid(dt, G, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"], f["sxz0"],
                                           f["eta_s"], x, z, ph, *out, interp_stencil=stencil)
        dp = ctx.update_particle_pressure_subgrid(dt, beta, f["p_in"], f["p0_n"], f["eta_n"], x, z, ph, P)
    for a, b in zip(out, exp_s):
        _close(a, b)
    _close(P, exp_p)
    assert np.array_equal(dp, exp_dp)
    on = ph != -1
    assert np.abs(out[0][on] - m["sxx"][on]).max() > 1e-3          # the update did act


def test_no_relaxation_equals_the_plain_update(oracle_mod):
    g, ctx, x, z, ph, f, m = _case(oracle_mod, 20, seed=3)
    plain = [m[k].copy() for k in ("sxx", "szz", "sxz")]
    ctx.update_particle_stress(f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"], f["sxz0"], x, z, ph, *plain)
    sub = [m[k].copy() for k in ("sxx", "szz", "sxz")]
    ctx.update_particle_stress_subgrid(1.0, [1e-300] * 3, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"],
                                       f["sxz0"], f["eta_s"], x, z, ph, *sub)
    for a, b in zip(plain, sub):
        assert np.array_equal(a, b)
    # markers that are not the ones set are refused
    with pytest.raises(mdgpu.MdgError):
        ctx.update_particle_stress_subgrid(1.0, [1.0] * 3, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"],
                                           f["sxz0"], f["eta_s"], x[:-1], z[:-1], ph[:-1], *[a[:-1] for a in sub])
