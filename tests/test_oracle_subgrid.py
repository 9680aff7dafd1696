"""Oracle pins of the subgrid-diffusion marker updates (UpdateParticleStress
with subgrid_diffusion = 2, RheologyParticles.c:1811-1895; UpdateParticle-
Pressure with subgrid_diffusion >= 2, :1621-1695), from what the
mathematics fixes:

* a vanishing relaxation (G -> 0 for stresses, beta -> 0 for the
  pressure, whose relaxation time is taken as eta / beta:
  exp(-dt / t_relax) = 1) leaves no subgrid part, so the update is
  the plain one bit for bit;
* markers that already carry the old grid field at their positions have no
  subgrid deviation, so the update is again the plain one, for any
  relaxation time;
* full relaxation (exp(...) = 0) of a marker deviation that is the same on
  every marker takes it out on the markers and puts it back through the
  grid unchanged (P2Mastah of a constant is that constant, its
  interpolation too), so with no grid change the markers keep their values:
  a wrong sign or a dropped remaining-increment term breaks this."""
import numpy as np
import pytest

from mdgpu import setups as S


def _case(O, n=20, seed=4):
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    rng = np.random.default_rng(seed)
    ph = rng.integers(0, 2, x.size).astype(np.int32)
    ph[::37] = -1
    nc, nv = n * n, (n + 1) ** 2
    f = {k: rng.standard_normal(nc) for k in ("sxxd", "szzd", "sxxd0", "szzd0", "p_in", "p0_n")}
    f.update({k: rng.standard_normal(nv) for k in ("sxz", "sxz0")})
    f["eta_s"] = rng.uniform(0.5, 2.0, nv)
    f["eta_n"] = rng.uniform(0.5, 2.0, nc)
    m = {k: rng.standard_normal(x.size) for k in ("sxx", "szz", "sxz", "P")}
    return g, x, z, ph, f, m


@pytest.mark.parametrize("stencil", [1, 9])
def test_no_relaxation_is_the_plain_update(oracle_mod, stencil):
    O = oracle_mod
    g, x, z, ph, f, m = _case(O)
    plain = O.update_particle_stress(g, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"], f["sxz0"], x, z, ph,
                                     m["sxx"], m["szz"], m["sxz"])
    sub = O.update_particle_stress_subgrid(g, 1.0, [1e-300, 1e-300], f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"],
                                           f["szzd0"], f["sxz0"], f["eta_s"], x, z, ph, m["sxx"], m["szz"], m["sxz"],
                                           stencil=stencil)
    for a, b in zip(plain, sub):
        assert np.array_equal(a, b)
    Pp, dpp = O.update_particle_pressure(g, f["p_in"], f["p0_n"], x, z, ph, m["P"])
    Ps, dps = O.update_particle_pressure_subgrid(g, 1.0, [1e-300, 1e-300], f["p_in"], f["p0_n"], f["eta_n"], x, z, ph,
                                                 m["P"])
    assert np.array_equal(Pp, Ps) and np.array_equal(dpp, dps)


def test_markers_on_the_old_field_have_no_subgrid_part(oracle_mod):
    O = oracle_mod
    g, x, z, ph, f, m = _case(O, seed=5)
    on = ph != -1
    t0 = [O.g2p_centroids(g, f["sxxd0"], x, z, ph), O.g2p_centroids(g, f["szzd0"], x, z, ph),
          O.g2p_vertices(g, f["sxz0"], x, z, ph)]
    plain = O.update_particle_stress(g, f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"], f["szzd0"], f["sxz0"], x, z, ph,
                                     *t0)
    sub = O.update_particle_stress_subgrid(g, 0.7, [2.0, 5.0], f["sxxd"], f["szzd"], f["sxz"], f["sxxd0"],
                                           f["szzd0"], f["sxz0"], f["eta_s"], x, z, ph, *t0)
    for a, b in zip(plain, sub):
        np.testing.assert_allclose(b[on], a[on], rtol=0, atol=1e-14)
    P0 = O.g2p_centroids(g, f["p0_n"], x, z, ph)
    Pp, _ = O.update_particle_pressure(g, f["p_in"], f["p0_n"], x, z, ph, P0)
    Ps, _ = O.update_particle_pressure_subgrid(g, 0.7, [0.3, 2.0], f["p_in"], f["p0_n"], f["eta_n"], x, z, ph, P0)
    np.testing.assert_allclose(Ps[on], Pp[on], rtol=0, atol=1e-14)


@pytest.mark.parametrize("stencil", [1, 9])
def test_full_relaxation_of_a_uniform_deviation(oracle_mod, stencil):
    O = oracle_mod
    g, x, z, ph, f, m = _case(O, seed=6)
    on = ph != -1
    c = {"sxx": 1.3, "szz": -0.4, "sxz": 0.25, "P": 2.0}
    nc, nv = f["sxxd"].size, f["sxz"].size
    grid = {k: np.full(nc if k != "sxz" else nv, v) for k, v in c.items()}
    delta = 0.37
    mk = {k: np.where(on, v + delta, 0.0) for k, v in c.items()}
    sub = O.update_particle_stress_subgrid(g, 1.0, [1e300, 1e300], grid["sxx"], grid["szz"], grid["sxz"],
                                           grid["sxx"], grid["szz"], grid["sxz"], f["eta_s"], x, z, ph, mk["sxx"],
                                           mk["szz"], mk["sxz"], stencil=stencil)
    for k, a in zip(("sxx", "szz", "sxz"), sub):
        np.testing.assert_allclose(a[on], mk[k][on], rtol=0, atol=1e-12)
    # a deviation that varies from marker to marker is smoothed, not kept
    rough = {k: np.where(on, c[k] + m[k], 0.0) for k in ("sxx", "szz", "sxz")}
    out = O.update_particle_stress_subgrid(g, 1.0, [1e300, 1e300], grid["sxx"], grid["szz"], grid["sxz"],
                                           grid["sxx"], grid["szz"], grid["sxz"], f["eta_s"], x, z, ph, rough["sxx"],
                                           rough["szz"], rough["sxz"], stencil=stencil)
    assert np.std(out[0][on] - c["sxx"]) < 0.8 * np.std(rough["sxx"][on] - c["sxx"])
    Ps, _ = O.update_particle_pressure_subgrid(g, 1.0, [1e300, 1e300], grid["P"], grid["P"], f["eta_n"], x, z, ph,
                                               mk["P"])
    np.testing.assert_allclose(Ps[on], mk["P"][on], rtol=0, atol=1e-12)
