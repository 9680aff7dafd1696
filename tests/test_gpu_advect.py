"""GPU V2P and RK4 marker advection (mdg_v2p, mdg_advect_rk4) against the
oracle's restatement of ParticleRoutines.c:63-256 and
AdvectionRoutines.c:320-410. Both are built without floating-point
contraction and follow the reference's operation order, so the comparison
is bit for bit: velocities, positions, phases. Cases: random velocity
fields, markers on and beyond the box edges, inactive markers, Taras'
correction on and off; and the
2047^2 workload's 37.7M markers, a subsample of which is checked against the
oracle while the whole set is timed. The air-node branches (tag 30) are not
reachable here: mdg_set_bc refuses free-surface tags, so they stay
unpinned."""
import time

import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S

pytestmark = pytest.mark.gpu


def _case(O, n, seed):
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    g.pure_shear(1.0)
    rng = np.random.default_rng(seed)
    u = rng.standard_normal(g.BCu_t.size)
    v = rng.standard_normal(g.BCv_t.size)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=seed)
    x = np.concatenate([x, [0.0, 1.0, -1e-3, 1.0 + 1e-3, 0.5, 0.5], rng.uniform(-0.02, 1.02, 500)])
    z = np.concatenate([z, [-0.5, -0.5, -0.5, -0.5, 0.0, -1.0], rng.uniform(-1.02, 0.02, 500)])
    ph = rng.integers(0, 3, x.size).astype(np.int32)
    ph[::53] = -1
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    return g, ctx, u, v, x, z, ph


@pytest.mark.parametrize("conserv", [0, 1])
@pytest.mark.parametrize("n", [63, 100])
def test_v2p_and_rk4_bit_exact(oracle_mod, conserv, n):
    O = oracle_mod
    g, ctx, u, v, x, z, ph = _case(O, n, 11 + conserv + n)
    vx, vz = ctx.v2p(u, v, x, z, ph, conserv=conserv)
    rvx, rvz = O.v2p(g, u, v, x, z, ph, conserv=conserv)
    assert np.array_equal(vx, rvx) and np.array_equal(vz, rvz)
    dt = 0.3 * g.dx   # a third of a cell per unit velocity
    xr, zr, phr, vxr, vzr = O.advect_rk4(g, u, v, dt, x, z, ph, conserv=conserv,
                                         vx=np.full(x.size, 9.0), vz=np.full(x.size, 9.0))
    xg, zg, phg = x.copy(), z.copy(), ph.copy()
    vxg, vzg = np.full(x.size, 9.0), np.full(x.size, 9.0)
    ctx.advect_rk4(u, v, dt, xg, zg, phg, conserv=conserv, vx=vxg, vz=vzg)
    print("left the box", int(((phr == -1) & (ph != -1)).sum()), "of", x.size, flush=True)
    assert np.array_equal(phg, phr)
    assert np.array_equal(xg, xr) and np.array_equal(zg, zr)
    assert np.array_equal(vxg, vxr) and np.array_equal(vzg, vzr)
    assert ((phr == -1) & (ph != -1)).any()


def test_rk4_full_size_2047(oracle_mod):
    """37.7M markers (9 per cell) advected by a smooth cellular flow on the
    2047^2 grid, on device arrays; a 1-in-97 subsample against the oracle."""
    O = oracle_mod
    from mdgpu import workloads as W
    n = 2047
    wl = W.ellipse_workload(n, 3, 2024, n_ell=4)
    x, z = wl["x"], wl["z"]
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    g.pure_shear(0.0)   # free-slip box (the flow below is tangential at the walls)
    xg, zvx, xvz, zg = (np.asarray(a) for a in (g.xg, g.zvx, g.xvz, g.zg))
    # stream function psi = sin(pi x) sin(pi z): u = dpsi/dz, v = -dpsi/dx
    u = (np.sin(np.pi * xg)[None, :] * np.pi * np.cos(np.pi * zvx)[:, None]).ravel()
    v = (-np.pi * np.cos(np.pi * xvz)[None, :] * np.sin(np.pi * zg)[:, None]).ravel()
    ph = np.zeros(x.size, np.int32)
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    dx_, dz_, dph = ctx.device(x.size, host=x), ctx.device(z.size, host=z), ctx.device(x.size, np.int32, ph)
    dvx, dvz = ctx.device(x.size, host=np.zeros(x.size)), ctx.device(x.size, host=np.zeros(x.size))
    du, dv = ctx.device(u.size, host=u), ctx.device(v.size, host=v)
    dt = 0.5 * g.dx / np.pi
    ms = []
    for rep in range(3):
        if rep:
            dx_.upload(x); dz_.upload(z); dph.upload(ph)
        ctx.advect_rk4(du, dv, dt, dx_, dz_, dph, vx=dvx, vz=dvz)
        ms.append(ctx.last_kernel_ms())
    k = ms[-1]
    print("markers %d, RK4 kernel %.3f ms, %.0f M markers/s, %.2f TB/s of marker streams (44 B/marker)"
          % (x.size, k, x.size / (k * 1e-3) / 1e6, 44.0 * x.size / (k * 1e-3) / 1e12), flush=True)
    sub = np.arange(0, x.size, 97)
    xr, zr, phr, vxr, vzr = O.advect_rk4(g, u, v, dt, x[sub], z[sub], ph[sub])
    assert np.array_equal(dx_.numpy()[sub], xr) and np.array_equal(dz_.numpy()[sub], zr)
    assert np.array_equal(dph.numpy()[sub], phr)
    assert np.array_equal(dvx.numpy()[sub], vxr) and np.array_equal(dvz.numpy()[sub], vzr)


@pytest.mark.parametrize("n", [63, 100])
def test_g2p_and_marker_updates_bit_exact(oracle_mod, n):
    """Interp_Grid2P_centroids2, Interp_Grid2P (vertices),
    UpdateParticlePressure and UpdateParticleStress on the GPU against
    the oracle: random centre fields, markers up to the walls and corners,
    inactive markers."""
    O = oracle_mod
    g, ctx, u, v, x, z, ph = _case(O, n, 5 + n)
    inside = (x >= 0.0) & (x <= 1.0) & (z >= -1.0) & (z <= 0.0)
    ph = np.where(inside, ph, -1).astype(np.int32)
    rng = np.random.default_rng(n)
    F = rng.standard_normal(n * n)
    assert np.array_equal(ctx.g2p_centroids(F, x, z, ph), O.g2p_centroids(g, F, x, z, ph))
    Fv = rng.standard_normal((n + 1) * (n + 1))
    assert np.array_equal(ctx.g2p_vertices(Fv, x, z, ph), O.g2p_vertices(g, Fv, x, z, ph))
    p0, p1, P0 = rng.standard_normal(n * n), rng.standard_normal(n * n), rng.standard_normal(x.size)
    Pr, dpr = O.update_particle_pressure(g, p1, p0, x, z, ph, P0)
    Pg, dpg = P0.copy(), np.zeros(n * n)
    ctx.update_particle_pressure(p1, p0, x, z, ph, Pg, dp=dpg)
    assert np.array_equal(dpg, dpr) and np.array_equal(Pg, Pr)
    nc, nv = n * n, (n + 1) * (n + 1)
    grid = [rng.standard_normal(m) for m in (nc, nc, nv, nc, nc, nv)]
    mk = [rng.standard_normal(x.size) for _ in range(3)]
    ref = O.update_particle_stress(g, *grid, x, z, ph, *mk)
    got = [a.copy() for a in mk]
    ctx.update_particle_stress(*grid, x, z, ph, *got)
    assert all(np.array_equal(a, b) for a, b in zip(got, ref))


def test_accumulated_strain_and_rotation_match_oracle(oracle_mod):
    """AccumulatedStrainII and the stress rotation on the GPU against the
    oracle: bit-exact for the strain accumulation and the Jaumann rate (same
    arithmetic, no contraction), 1e-15 relative for the analytical rotation
    (device cos/sin against glibc's); the reference's own sequence (the spin
    P2Mastah of a zero marker array, then rotation with wxzm = 0) leaves the
    stresses unchanged and zeroes the vertex spin."""
    O = oracle_mod
    n = 41
    g = O.Grid(n, n + 6, 0.0, 1.0, -1.3, 0.0)
    g.pure_shear(1.0)
    ctx = mdgpu.Context(n, n + 6, 0.0, 1.0, -1.3, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=7)
    ph = np.zeros(x.size, np.int32)
    ph[::13] = -1
    rng = np.random.default_rng(3)
    nc = (n - 1) * (n + 5)
    rates = {k: np.abs(rng.standard_normal(nc)) for k in O.STRAIN_RATES[:6]}
    strains = {k: rng.standard_normal(x.size) for k in O.STRAINS}
    divpl = rng.standard_normal(nc)
    ref = O.accumulated_strain(g, 0.7, x, z, ph, rates, strains, div_u_pl=divpl)
    got = {k: v.copy() for k, v in strains.items()}
    ctx.accumulated_strain(0.7, x, z, ph, rates, got, div_u_pl=divpl)
    for k in O.STRAINS:
        assert np.array_equal(got[k], ref[k]), k
    with pytest.raises(mdgpu.MdgError):
        ctx.accumulated_strain(0.7, x, z, ph, {"eII_pl": -np.ones(nc)}, {"strain": np.zeros(x.size)})
    txx, tzz, txz, w = rng.standard_normal((4, x.size))
    for mode in (1, 2):
        r = O.rotate_stress(mode, 0.2, ph, w, txx, tzz, txz)
        s = [a.copy() for a in (txx, tzz, txz)]
        ctx.rotate_particle_stress(mode, 0.2, ph, w, *s)
        for a, b in zip(s, r):
            if mode == 1:
                assert np.array_equal(a, b)
            else:
                np.testing.assert_allclose(a, b, rtol=1e-15, atol=1e-15 * np.abs(b).max())
    # the reference's call sequence: P2Mastah of the zero marker spin onto
    # the vertices (flag 1), then the rotation with that zero spin
    wxz = rng.standard_normal(n * (n + 6))
    wxzm = np.zeros(x.size)
    ctx.markers_set(x, z, ph)
    ctx.p2mastah(wxzm, node_field=wxz, node_type=g.BCg_t, flag=1, avg=0, centroid=0)
    assert np.all(wxz == 0.0)
    s = [a.copy() for a in (txx, tzz, txz)]
    ctx.rotate_particle_stress(2, 0.2, ph, wxzm, *s)
    for a, b in zip(s, (txx, tzz, txz)):
        assert np.array_equal(a, b)
