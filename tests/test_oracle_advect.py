"""Oracle V2P and RK4 advection (ParticleRoutines.c:63-256,
AdvectionRoutines.c:320-410) pinned by what the mathematics fixes: bilinear
V2P reproduces linear velocity fields at every marker of the box (with or
without Taras' correction, whose second differences vanish on them); the
classic RK4 step on a linear autonomous field is the fourth-order Taylor
polynomial of exp(dt A); markers that leave the box are flagged inactive and
stay where they left; inactive markers do not move."""
import numpy as np
import pytest

from mdgpu import setups as S


def _linear(g, a):
    """u, v on the Vx / Vz nodes of u = a0 + a1 x + a2 z, v = a3 + a4 x + a5 z."""
    xg, zvx = np.asarray(g.xg), np.asarray(g.zvx)
    xvz, zg = np.asarray(g.xvz), np.asarray(g.zg)
    u = (a[0] + a[1] * xg[None, :] + a[2] * zvx[:, None]).ravel()
    v = (a[3] + a[4] * xvz[None, :] + a[5] * zg[:, None]).ravel()
    return u, v


@pytest.mark.parametrize("conserv", [0, 1])
def test_v2p_reproduces_linear_fields(oracle_mod, conserv):
    O = oracle_mod
    g = O.Grid(41, 33, -1.0, 1.0, 0.0, 0.8)
    a = (0.3, 0.5, -0.2, -0.1, 0.25, -0.5)
    u, v = _linear(g, a)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=4)
    ph = np.zeros(x.size, np.int32)
    ph[::97] = -1
    vx, vz = O.v2p(g, u, v, x, z, ph, conserv=conserv)
    act = ph != -1
    np.testing.assert_allclose(vx[act], (a[0] + a[1] * x + a[2] * z)[act], rtol=0, atol=2e-15)
    np.testing.assert_allclose(vz[act], (a[3] + a[4] * x + a[5] * z)[act], rtol=0, atol=2e-15)
    assert (vx[~act] == 0).all() and (vz[~act] == 0).all()


def test_rk4_on_a_linear_field_is_the_taylor_polynomial(oracle_mod):
    O = oracle_mod
    g = O.Grid(65, 65, -1.0, 1.0, -1.0, 1.0)
    # rotation plus pure shear about the origin: d/dt (x, z) = A (x, z)
    A = np.array([[0.2, -1.0], [0.7, -0.2]])
    u, v = _linear(g, (0.0, A[0, 0], A[0, 1], 0.0, A[1, 0], A[1, 1]))
    rng = np.random.default_rng(2)
    x, z = rng.uniform(-0.4, 0.4, 2000), rng.uniform(-0.4, 0.4, 2000)
    dt = 0.05
    xn, zn, ph, vx, vz = O.advect_rk4(g, u, v, dt, x, z, np.zeros(x.size, np.int32))
    hA = dt * A
    T = np.eye(2) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    ref = T @ np.stack([x, z])
    assert (ph == 0).all()
    np.testing.assert_allclose(xn, ref[0], rtol=0, atol=1e-15)
    np.testing.assert_allclose(zn, ref[1], rtol=0, atol=1e-15)
    # the stored velocity is the RK4 average: (x_new - x) / dt
    np.testing.assert_allclose(vx, (xn - x) / dt, rtol=0, atol=1e-13)
    np.testing.assert_allclose(vz, (zn - z) / dt, rtol=0, atol=1e-13)


def test_markers_leaving_the_box_become_inactive(oracle_mod):
    O = oracle_mod
    g = O.Grid(21, 21, 0.0, 1.0, 0.0, 1.0)
    u, v = _linear(g, (1.0, 0.0, 0.0, 0.0, 0.0, 0.0))   # uniform flow in +x
    x = np.array([0.5, 0.97, 0.99, 0.2])
    z = np.array([0.5, 0.5, 0.5, 0.5])
    ph = np.array([0, 0, 0, -1], np.int32)
    vx0 = np.full(4, 7.0)
    xn, zn, phn, vx, vz = O.advect_rk4(g, u, v, 0.05, x, z, ph, vx=vx0, vz=vx0)
    assert list(phn) == [0, -1, -1, -1]
    assert xn[0] == pytest.approx(0.55, abs=1e-15)
    # 0.99 leaves at the first half step: stays at 0.99 + 0.025
    assert xn[2] == pytest.approx(1.015, abs=1e-15)
    # 0.97 leaves at the final step (0.97 + 0.05 = 1.02 > 1)
    assert xn[1] == pytest.approx(1.02, abs=1e-15)
    assert xn[3] == 0.2 and zn[3] == 0.5
    # velocities are written only for markers still active at the end
    assert vx[0] == pytest.approx(1.0) and vx[1] == 7.0 and vx[2] == 7.0 and vx[3] == 7.0


def _centre_linear(g, a):
    xc, zc = np.asarray(g.xc), np.asarray(g.zc)
    return (a[0] + a[1] * xc[None, :] + a[2] * zc[:, None]).ravel()


def test_g2p_centroids_linear_constant_and_edges(oracle_mod):
    """Interp_Grid2P_centroids2: a constant centre field is returned at every
    marker to round-off (the weights are normalised); a linear one is reproduced at
    the markers between the outermost centres; within half a cell of a wall
    the value is copied along the wall (linear in the tangential coordinate
    only); inactive markers get 0."""
    O = oracle_mod
    g = O.Grid(33, 25, 0.0, 2.0, -1.0, 0.5)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=8)
    ph = np.zeros(x.size, np.int32)
    ph[::89] = -1
    act = ph != -1
    out = O.g2p_centroids(g, np.full((g.Nx - 1) * (g.Nz - 1), 3.25), x, z, ph)
    np.testing.assert_allclose(out[act], 3.25, rtol=2e-15, atol=0)
    assert (out[~act] == 0).all()
    a = (0.5, 1.5, -2.0)
    out = O.g2p_centroids(g, _centre_linear(g, a), x, z, ph)
    xc, zc = np.asarray(g.xc), np.asarray(g.zc)
    inner = act & (x > xc[0]) & (x < xc[-1]) & (z > zc[0]) & (z < zc[-1])
    np.testing.assert_allclose(out[inner], (a[0] + a[1] * x + a[2] * z)[inner], rtol=0, atol=5e-15)
    west = act & (x < xc[0]) & (z > zc[0]) & (z < zc[-1])
    np.testing.assert_allclose(out[west], (a[0] + a[1] * xc[0] + a[2] * z)[west], rtol=0, atol=5e-15)
    corner = act & (x < xc[0]) & (z < zc[0])
    np.testing.assert_allclose(out[corner], a[0] + a[1] * xc[0] + a[2] * zc[0], rtol=0, atol=5e-15)
    assert inner.any() and west.any() and corner.any()


def test_update_particle_pressure_adds_the_interpolated_increment(oracle_mod):
    O = oracle_mod
    g = O.Grid(21, 21, 0.0, 1.0, 0.0, 1.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 2, 2, seed=3)
    ph = np.zeros(x.size, np.int32)
    rng = np.random.default_rng(0)
    p0 = rng.standard_normal(400)
    p1 = p0 + _centre_linear(g, (0.1, 0.2, -0.3))
    P0 = rng.standard_normal(x.size)
    P, dp = O.update_particle_pressure(g, p1, p0, x, z, ph, P0)
    np.testing.assert_allclose(dp, p1 - p0, rtol=0, atol=0)
    np.testing.assert_allclose(P, P0 + O.g2p_centroids(g, dp, x, z, ph), rtol=0, atol=0)


def test_g2p_vertices_reproduces_bilinear_fields(oracle_mod):
    """Interp_Grid2P on the vertices: the weights of the enclosing cell sum to
    one, so bilinear vertex fields (a + b x + c z + d x z) are reproduced at
    every marker without the normalisation the reference leaves out."""
    O = oracle_mod
    g = O.Grid(29, 37, -0.5, 0.5, -1.0, 0.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=6)
    ph = np.zeros(x.size, np.int32)
    ph[::71] = -1
    act = ph != -1
    xg, zg = np.asarray(g.xg), np.asarray(g.zg)
    a = (0.2, -1.0, 0.7, 1.3)
    F = (a[0] + a[1] * xg[None, :] + a[2] * zg[:, None] + a[3] * xg[None, :] * zg[:, None]).ravel()
    out = O.g2p_vertices(g, F, x, z, ph)
    np.testing.assert_allclose(out[act], (a[0] + a[1] * x + a[2] * z + a[3] * x * z)[act], rtol=0, atol=5e-15)
    assert (out[~act] == 0).all()


def test_update_particle_stress_adds_the_interpolated_increments(oracle_mod):
    """UpdateParticleStress without subgrid diffusion or rotation: each marker
    stress grows by the interpolated grid increment (centres for the normal
    components, vertices for the shear one); constant increments are added
    exactly as constants."""
    O = oracle_mod
    g = O.Grid(17, 17, 0.0, 1.0, 0.0, 1.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 2, 2, seed=9)
    ph = np.zeros(x.size, np.int32)
    nc, nv = 16 * 16, 17 * 17
    rng = np.random.default_rng(4)
    s0 = [rng.standard_normal(m) for m in (nc, nc, nv)]
    s1 = [s0[0] + 0.5, s0[1] - 0.25, s0[2] + 2.0]
    m0 = [rng.standard_normal(x.size) for _ in range(3)]
    mxx, mzz, mxz = O.update_particle_stress(g, *s1, *s0, x, z, ph, *m0)
    np.testing.assert_allclose(mxx, m0[0] + 0.5, rtol=0, atol=4e-15)
    np.testing.assert_allclose(mzz, m0[1] - 0.25, rtol=0, atol=4e-15)
    np.testing.assert_allclose(mxz, m0[2] + 2.0, rtol=0, atol=8e-15)
    dxx = s1[0] - s0[0]
    np.testing.assert_allclose(mxx, m0[0] + O.g2p_centroids(g, dxx, x, z, ph), rtol=0, atol=0)
