"""Oracle pins of the air-node (tag 30/31) branches of the grid->marker
interpolations, on hand-built tag arrays with a band of air along the top of
the box (the free-surface layout).  `mdg_set_bc` refuses air tags, so these
branches run on the CPU oracle only; the answers follow from fields that
vary along one axis:

* V2P (ParticleRoutines.c:99-132 for Vx, :176-200 for Vz): with a field
  linear in x, vertical substitutions change nothing, so markers below the
  band and in its first cell row see the field exactly; a cell whose two
  rows are air collapses to one column: the west one for Vx, the EAST one for
  Vz, since the reference fills an air Vz south-west node from the
  north-east one (:190) - "fixing" that index to north-west gives the west
  column and fails here.  With a field linear in z, a cell with an air north
  row sees its south row, and a cell with two air rows its north row.
* Interp_Grid2P_centroids2 / Interp_Grid2P (:1332-1580, :1223-1325): air
  nodes take their vertical neighbour, so a field linear in x is exact
  everywhere and a field linear in z gives the south row's value in a cell
  with an air north row.
* UpdateParticlePressure (RheologyParticles.c:1631-1637): the increment is
  zero on air cells, and a unit increment elsewhere reaches the markers
  below the band unchanged."""
import numpy as np

from mdgpu import setups as S

ZA = -0.3          # nodes above this height are air


def _grid(O, n=16):
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    Nx, Nz = g.Nx, g.Nz
    g.BCu_t[:] = np.where(np.repeat(g.zvx, Nx) > ZA, 30, -1)
    g.BCv_t[:] = np.where(np.repeat(g.zg, Nx + 1) > ZA, 30, -1)
    g.BCp_t[:] = np.where(np.repeat(g.zc, Nx - 1) > ZA, 31, -1)
    g.BCg_t[:] = np.where(np.repeat(g.zg, Nx) > ZA, 30, -1)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=7)
    return g, x, z, np.zeros(x.size, np.int32)


def _cell(c0, h, n, p):
    # the reference's ceil(d/h) - 1, clamped to the n - 1 intervals of n nodes
    return np.clip(np.ceil((p - c0) / h).astype(np.int64) - 1, 0, n - 2)


def test_v2p_air_band(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _grid(O)
    Nx, Nz = g.Nx, g.Nz
    a, b = 0.3, 1.7
    ju, iu = _cell(g.xg[0], g.dx, Nx, x), _cell(g.zvx[0], g.dz, Nz + 1, z)
    jv, iv = _cell(g.xvz[0], g.dx, Nx + 1, x), _cell(g.zg[0], g.dz, Nz, z)
    air_u = (g.zvx[iu] > ZA, g.zvx[iu + 1] > ZA)
    air_v = (g.zg[iv] > ZA, g.zg[iv + 1] > ZA)
    assert air_u[1].any() and (air_u[1] & ~air_u[0]).any() and (air_u[0] & air_u[1]).any()
    assert (air_v[0] & air_v[1]).any() and (~air_v[0] & air_v[1]).any()

    # linear in x
    u = a + b * np.tile(g.xg, Nz + 1)
    v = a + b * np.tile(g.xvz, Nz)
    vx, vz = O.v2p(g, u, v, x, z, ph)
    both_u, both_v = air_u[0] & air_u[1], air_v[0] & air_v[1]
    np.testing.assert_allclose(vx[~both_u], a + b * x[~both_u], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vx[both_u], a + b * g.xg[ju[both_u]], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vz[~both_v], a + b * x[~both_v], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vz[both_v], a + b * g.xvz[jv[both_v] + 1], rtol=0, atol=1e-14)
    # with Taras' correction, cells whose rows and x-neighbours are all below the band
    vx1, vz1 = O.v2p(g, u, v, x, z, ph, conserv=1)
    np.testing.assert_allclose(vx1[~air_u[1]], a + b * x[~air_u[1]], rtol=0, atol=1e-14)

    # linear in z
    u = a + b * np.repeat(g.zvx, Nx)
    v = a + b * np.repeat(g.zg, Nx + 1)
    vx, vz = O.v2p(g, u, v, x, z, ph)
    north_u, north_v = air_u[1] & ~air_u[0], air_v[1] & ~air_v[0]
    clear_u, clear_v = ~air_u[1], ~air_v[1]
    np.testing.assert_allclose(vx[clear_u], a + b * z[clear_u], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vx[north_u], a + b * g.zvx[iu[north_u]], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vx[both_u], a + b * g.zvx[iu[both_u] + 1], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vz[clear_v], a + b * z[clear_v], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vz[north_v], a + b * g.zg[iv[north_v]], rtol=0, atol=1e-14)
    np.testing.assert_allclose(vz[both_v], a + b * g.zg[iv[both_v] + 1], rtol=0, atol=1e-14)


def test_g2p_air_band(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _grid(O)
    Nx, Nz = g.Nx, g.Nz
    a, b = -0.8, 2.3
    # centres: the interpolation cell spans xc[j-1], xc[j] and zc[i-1], zc[i]
    # (extended grids xvz, zvx); the half cells along the walls copy the
    # edge values, so only the inner cells interpolate
    jc = np.ceil((x - g.xvz[0]) / g.dx).astype(np.int64) - 1
    ic = np.ceil((z - g.zvx[0]) / g.dz).astype(np.int64) - 1
    inner = (jc > 0) & (jc < Nx - 1) & (ic > 0) & (ic < Nz - 1)
    # linear in x: exact at every (inner, for the centres) marker, air rows or not
    Fc = a + b * np.tile(g.xc, Nz - 1)
    Fv = a + b * np.tile(g.xg, Nz)
    out = O.g2p_centroids(g, Fc, x, z, ph)
    np.testing.assert_allclose(out[inner], a + b * x[inner], rtol=0, atol=1e-13)
    np.testing.assert_allclose(O.g2p_vertices(g, Fv, x, z, ph), a + b * x, rtol=0, atol=1e-13)
    # linear in z: the first cell row under the band sees its south row
    Fc = a + b * np.repeat(g.zc, Nx - 1)
    Fv = a + b * np.repeat(g.zg, Nx)
    iv = _cell(g.zg[0], g.dz, Nz, z)
    north = (g.zg[iv + 1] > ZA) & ~(g.zg[iv] > ZA)
    clear = ~(g.zg[iv + 1] > ZA)
    out = O.g2p_vertices(g, Fv, x, z, ph)
    np.testing.assert_allclose(out[clear], a + b * z[clear], rtol=0, atol=1e-13)
    np.testing.assert_allclose(out[north], a + b * g.zg[iv[north]], rtol=0, atol=1e-13)
    zs, zn = g.zc[np.clip(ic - 1, 0, Nz - 2)], g.zc[np.clip(ic, 0, Nz - 2)]
    north = inner & (zn > ZA) & ~(zs > ZA)
    clear = inner & ~(zn > ZA)
    out = O.g2p_centroids(g, Fc, x, z, ph)
    assert north.any() and clear.any()
    np.testing.assert_allclose(out[clear], a + b * z[clear], rtol=0, atol=1e-13)
    np.testing.assert_allclose(out[north], a + b * zs[north], rtol=0, atol=1e-13)


def test_pressure_increment_skips_air(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _grid(O)
    nc = (g.Nx - 1) * (g.Nz - 1)
    g.BCp_t[:] = np.where(g.BCp_t == 31, 30, -1)          # tag 30 cells: zero increment
    P, dp = O.update_particle_pressure(g, np.full(nc, 3.0), np.full(nc, 2.0), x, z, ph, np.zeros(x.size))
    air = g.BCp_t == 30
    assert air.any() and np.all(dp[air] == 0.0) and np.all(dp[~air] == 1.0)
    ic = np.ceil((z - g.zvx[0]) / g.dz).astype(np.int64) - 1
    below = g.zc[np.clip(ic, 0, g.Nz - 2)] <= ZA
    np.testing.assert_allclose(P[below], 1.0, rtol=0, atol=1e-15)
    assert np.all((P >= -1e-15) & (P <= 1.0 + 1e-15))
