"""Pins of the oracle rheology (ViscosityConcise / NonNewtonianViscosityGrid,
RheologyDensity.c:152-197, 385-1080, 1131-1426) against closed forms and an
independent scalar root solve.  CPU only."""
import math

import numpy as np
import pytest
from scipy.optimize import brentq

import oracle as O

EXX, EZZ, EXZ = 0.7, -0.4, 0.25
EII = math.sqrt(0.5 * (EXX ** 2 + EZZ ** 2 + (EXX + EZZ) ** 2) + EXZ ** 2)


def vc(ph, md, G=1.0, T=1.0, P=0.0, **kw):
    return O.viscosity_concise(ph, md, G, T, P, EXX, EZZ, EXZ, **kw)


def test_invariant_uses_out_of_plane_component():
    # Eyy = -(Exx + Ezz) enters the second invariant (RheologyDensity.c:457-458)
    ph = O.phase(cstv=0, pwlv=1, n_pwl=3.0)
    eta, etaVE, *_ = vc(ph, O.rheo_model())
    assert etaVE == pytest.approx(EII ** (1 / 3 - 1), rel=1e-14)


def test_constant_viscosity():
    eta, etaVE, pc, epl, st = vc(O.phase(eta0=3.5), O.rheo_model(), P=2.0)
    assert (eta, etaVE, pc, epl, st) == (3.5, 3.5, 2.0, 0.0, 0)


def test_maxwell_viscoelastic():
    G, dt, eta0 = 2.0, 0.3, 1.7
    md = O.rheo_model(elastic=1, dt=dt)
    eta, etaVE, *_ = vc(O.phase(eta0=eta0), md, G=G)
    assert etaVE == pytest.approx(1.0 / (1.0 / eta0 + 1.0 / (G * dt)), rel=1e-13)
    assert eta == pytest.approx(eta0, rel=1e-14)   # dissipative part excludes the spring
    # step 0 switches elasticity off (RheologyDensity.c:483)
    _, etaVE0, *_ = vc(O.phase(eta0=eta0), O.rheo_model(elastic=1, dt=dt, step=0), G=G)
    assert etaVE0 == eta0


@pytest.mark.parametrize("n", [1.0, 2.5, 3.5])
def test_power_law_arrhenius(n):
    A, F, pref, Q, V, R, T, P = 3.0, 0.6, 1.3, 2.0, 0.5, 1.1, 0.9, 0.4
    ph = O.phase(cstv=0, pwlv=1, n_pwl=n, A_pwl=A, F_pwl=F, pref_pwl=pref, Q_pwl=Q, V_pwl=V)
    md = O.rheo_model(R=R)
    eta, etaVE, *_ = vc(ph, md, T=T, P=P)
    B = pref * F * A ** (-1 / n) * math.exp((Q + P * V) / (n * R * T))
    assert etaVE == pytest.approx(B * EII ** (1 / n - 1), rel=1e-11)
    assert eta == pytest.approx(etaVE, rel=1e-10)
    # zeroC floor on the temperature (RheologyDensity.c:496)
    _, e_lo, *_ = vc(ph, O.rheo_model(R=R, zeroC=T), T=0.1, P=P)
    assert e_lo == pytest.approx(etaVE, rel=1e-12)
    # pref multiplier scales the prefactor (per-node extension; 1 = reference)
    _, e2, *_ = vc(ph, md, T=T, P=P, pref_mult=2.0)
    assert e2 == pytest.approx(2.0 * etaVE, rel=1e-11)


def test_local_newton_matches_independent_root():
    """Constant + power law + Maxwell: the stress solves the strain-rate
    partition E = T/(2 G dt) + T/(2 eta0) + (2B)^-n T^n (bracketed root)."""
    G, dt, eta0, n, A = 1.5, 0.4, 2.0, 3.0, 0.8
    ph = O.phase(cstv=1, eta0=eta0, pwlv=1, n_pwl=n, A_pwl=A)
    md = O.rheo_model(elastic=1, dt=dt)
    eta, etaVE, pc, epl, st = vc(ph, md, G=G)
    B = A ** (-1 / n)
    f = lambda T: T / (2 * G * dt) + T / (2 * eta0) + (2 * B) ** (-n) * T ** n - EII
    T = brentq(f, 0.0, 100.0, xtol=1e-15, rtol=1e-15)
    assert st == 0
    assert etaVE == pytest.approx(T / (2 * EII), rel=1e-10)
    e_pwl, e_cst = (2 * B) ** (-n) * T ** n, T / (2 * eta0)
    assert eta == pytest.approx(T / (2 * (e_pwl + e_cst)), rel=1e-9)


def _trial(eta0, G, dt):
    return 1.0 / (1.0 / eta0 + 1.0 / (G * dt))


def test_drucker_prager_rate_independent():
    G, dt, eta0, C, phi, P = 1.0, 1.0, 10.0, 0.3, math.radians(30), 0.5
    ph = O.phase(eta0=eta0, plast=1, C=C, fric=phi)
    md = O.rheo_model(elastic=1, dt=dt)
    eta, etaVE, pc, epl, st = vc(ph, md, G=G, P=P)
    ty = C * math.cos(phi) + P * math.sin(phi)
    eve = _trial(eta0, G, dt)
    assert 2 * eve * EII > ty
    assert etaVE == pytest.approx(ty / (2 * EII), rel=1e-12)
    gdot = (2 * eve * EII - ty) / eve
    assert epl == pytest.approx(gdot / 2, rel=1e-10)
    assert eta == pytest.approx(1.0 / (1.0 / eta0 + gdot / ty), rel=1e-10)
    assert pc == P   # no dilation


def test_drucker_prager_below_yield_is_viscoelastic():
    ph = O.phase(eta0=0.1, plast=1, C=10.0, fric=0.5)
    md = O.rheo_model(elastic=1, dt=1.0)
    _, etaVE, _, epl, _ = vc(ph, md)
    assert etaVE == pytest.approx(_trial(0.1, 1.0, 1.0), rel=1e-13) and epl == 0.0
    # step 0 switches plasticity off as well
    _, e0, *_ = vc(O.phase(eta0=10.0, plast=1, C=0.01), O.rheo_model(step=0, elastic=1))
    assert e0 == 10.0


def test_viscoplastic_regularisation_linear():
    """n_vp = 1: one return-map step is exact, tau = ty + eta_vp*F/(eta_ve+eta_vp)."""
    eta0, C, eta_vp = 5.0, 0.2, 0.7
    ph = O.phase(eta0=eta0, plast=1, C=C, fric=0.0, eta_vp=eta_vp, n_vp=1.0)
    md = O.rheo_model(elastic=1, dt=1.0)
    _, etaVE, *_ = vc(ph, md)
    eve = _trial(eta0, 1.0, 1.0)
    F = 2 * eve * EII - C
    assert etaVE == pytest.approx((C + eta_vp * F / (eve + eta_vp)) / (2 * EII), rel=1e-12)


def test_von_mises_cap_and_tension_floor():
    ph = O.phase(eta0=5.0, plast=1, C=1.0, fric=0.4, Slim=0.05)
    md = O.rheo_model(elastic=1, dt=1.0)
    _, etaVE, *_ = vc(ph, md, P=3.0)
    assert etaVE == pytest.approx(0.05 / (2 * EII), rel=1e-12)
    # yield stress < 0 in tension: the return map ends at Tiimin (:694-701)
    ph = O.phase(eta0=5.0, plast=1, C=0.0, fric=0.5)
    md = O.rheo_model(elastic=1, dt=1.0, Tiimin=1e-3)
    _, etaVE, pc, epl, _ = vc(ph, md, P=-4.0)
    assert etaVE == pytest.approx(1e-3 / (2 * EII), rel=1e-12) and pc == -4.0
    eve = _trial(5.0, 1.0, 1.0)
    assert epl == pytest.approx((2 * eve * EII - 1e-3) / eve / 2, rel=1e-12)


def test_dilatant_pressure_correction():
    """Dilation: Pc = P + K dt sin(psi) gdot with gdot -> F/eta_ve."""
    eta0, C, phi, psi, beta, dt = 5.0, 0.2, 0.5, 0.2, 10.0, 1.0   # K = 0.1
    ph = O.phase(eta0=eta0, plast=1, C=C, fric=phi, dil=psi, beta_for_K=beta)
    md = O.rheo_model(elastic=1, dt=dt)
    P = 0.3
    _, etaVE, pc, epl, _ = vc(ph, md, P=P)
    eve = _trial(eta0, 1.0, dt)
    ty = C * math.cos(phi) + P * math.sin(phi)
    gdot = (2 * eve * EII - ty) / eve
    assert etaVE == pytest.approx(ty / (2 * EII), rel=1e-10)
    assert pc == pytest.approx(P + dt / beta * math.sin(psi) * gdot, rel=1e-10)


def test_clipping():
    md = O.rheo_model(min_eta=0.5, max_eta=2.0)
    assert vc(O.phase(eta0=10.0), md)[:2] == (2.0, 2.0)
    assert vc(O.phase(eta0=0.01), md)[:2] == (0.5, 0.5)


@pytest.mark.parametrize("average", [0, 1, 2])
def test_grid_phase_averages_and_stresses(average):
    rng = np.random.default_rng(3)
    n = 50
    f = rng.uniform(0, 1, n)
    f[:5] = 0.0
    f[5:8] = 1e-14            # below the 1e-13 activity threshold
    perc = np.stack([1 - f, f])
    eta1, eta2 = 2.0, 1e6     # the stiff phase must vanish where inactive
    ph = [O.phase(eta0=eta1), O.phase(eta0=eta2)]
    G = rng.uniform(1, 2, n)
    dt = 0.5
    md = O.rheo_model(elastic=1, dt=dt)
    exx, ezz, exz = rng.normal(size=(3, n))
    s0 = rng.normal(size=(3, n))
    out = O.rheology_nodes(md, ph, perc, G, np.ones(n), np.zeros(n), exx, ezz, exz,
                           sxx0=s0[0], szz0=s0[1], sxz0=s0[2], average=average)
    ve = lambda e: 1 / (1 / e + 1 / (G * dt))
    a1, a2 = ve(eta1), ve(eta2)
    act2 = f > 1e-13
    w1, w2 = 1 - f, np.where(act2, f, 0.0)
    if average == 0:
        ref, refp = w1 * a1 + w2 * a2, w1 * eta1 + w2 * eta2
    elif average == 1:
        ref, refp = 1 / (w1 / a1 + w2 / a2), 1 / (w1 / eta1 + w2 / eta2)
    else:
        ref = np.exp(w1 * np.log(a1) + w2 * np.log(a2))
        refp = np.exp(w1 * np.log(eta1) + w2 * np.log(eta2))
    np.testing.assert_allclose(out["eta"], ref, rtol=1e-12)
    np.testing.assert_allclose(out["eta_phys"], refp, rtol=1e-12)
    Exx = exx + s0[0] / (G * dt) / 2
    Exz = exz + s0[2] / (G * dt) / 2
    np.testing.assert_allclose(out["sxx"], 2 * ref * Exx, rtol=1e-12)
    np.testing.assert_allclose(out["sxz"], 2 * ref * Exz, rtol=1e-12)


def test_grid_air_nodes_are_zero():
    n = 6
    tag = np.array([-1, 30, 31, -1, 30, 31], np.int8)
    kw = dict(G=np.ones(n), T=np.ones(n), P=np.zeros(n), exx=np.ones(n), ezz=-np.ones(n),
              exz=np.ones(n), tag=tag)
    perc = np.ones((1, n))
    c = O.rheology_nodes(O.rheo_model(), [O.phase(eta0=2.0)], perc, centroid=1, **kw)
    v = O.rheology_nodes(O.rheo_model(), [O.phase(eta0=2.0)], perc, centroid=0, **kw)
    np.testing.assert_array_equal(c["eta"], [2, 0, 0, 2, 0, 0])     # 30 and 31 on centres
    np.testing.assert_array_equal(v["eta"], [2, 0, 2, 2, 0, 2])     # only 30 on vertices
    np.testing.assert_array_equal(c["sxx"], [4, 0, 0, 4, 0, 0])


def test_exz_n_reproduces_linear_field():
    g = O.Grid(9, 7, 0.0, 1.0, 0.0, 0.7)
    g.pure_shear(0.0)
    X, Z = np.meshgrid(g.xg, g.zg)
    exz = (0.3 * X - 1.2 * Z + 0.1).ravel()
    xc, zc = np.meshgrid(g.xc, g.zc)
    np.testing.assert_allclose(O.exz_n(g, exz), (0.3 * xc - 1.2 * zc + 0.1).ravel(), atol=1e-14)


def _soft(**kw):
    d = dict(G=1.0, bet=1.0, C=1.0, C_end=0.2, phi=0.5, phi_end=0.1, psi=0.2, psi_end=0.0,
             pls_start=0.1, pls_end=1.1, H_c=0.0, phi_soft=1, psi_soft=1, coh_soft=1)
    d.update(kw)
    return O.OrcPhaseSoft(**d)


def test_material_grid_harmonic_moduli():
    perc = np.array([[0.25, 1.0, 0.0], [0.75, 0.0, 1.0]])
    out = O.material_grid([_soft(G=2.0, bet=4.0), _soft(G=8.0, bet=1.0)], perc)
    np.testing.assert_allclose(out["mu"], [1 / (0.25 / 2 + 0.75 / 8), 2.0, 8.0], rtol=1e-15)
    np.testing.assert_allclose(out["bet"], [1 / (0.25 / 4 + 0.75), 4.0, 1.0], rtol=1e-15)


def test_linear_softening_closed_form():
    """style 0: piecewise linear in the accumulated strain, with the
    reference's slope times e (RheologyDensity.c:1478-1497)."""
    e = np.array([0.0, 0.05, 0.1, 0.6, 1.1, 3.0])
    out = O.material_grid([_soft()], np.ones((1, e.size)), strain_pl=e, style=0)
    slope = (0.2 - 1.0) / 1.0
    want = np.where(e < 0.1, 1.0, np.where(e >= 1.1, 0.2, 1.0 + slope * e))
    np.testing.assert_allclose(out["C"], want, rtol=1e-15)
    np.testing.assert_allclose(out["fric"], np.where(e < 0.1, 0.5, np.where(e >= 1.1, 0.1, 0.5 - 0.4 * e)),
                               rtol=1e-15)
    # no softening flag keeps the initial value
    out = O.material_grid([_soft(coh_soft=0)], np.ones((1, e.size)), strain_pl=e, style=0)
    np.testing.assert_array_equal(out["C"], 1.0)


def test_erfc_softening_limits_and_hardening():
    from math import erfc
    e = np.array([0.0, 0.6, 50.0])
    out = O.material_grid([_soft()], np.ones((1, 3)), strain_pl=e, style=1)
    # exactly the initial value at zero strain, the end value far beyond pls_end
    assert out["C"][0] == pytest.approx(1.0, rel=1e-14)
    assert out["C"][2] == pytest.approx(0.2, rel=1e-12)
    mu, ds, d = 0.6, 1.0, 0.8
    f = lambda x: 1.0 - d / 2 * erfc(-(x - mu) / ds) - 0.2
    assert out["C"][1] == pytest.approx(f(0.6) * d / f(0.0) + 0.2, rel=1e-14)
    out = O.material_grid([_soft(H_c=-0.5)], np.ones((1, 3)), strain_pl=e, style=2)
    np.testing.assert_allclose(out["C"], np.maximum(1.0 - 0.5 * e, 0.2), rtol=1e-15)


def test_tangent_constant_viscosity():
    n = 4
    D = O.tangent_nodes(O.rheo_model(), [O.phase(eta0=3.0)], np.ones((1, n)), np.ones(n),
                        np.ones(n), np.zeros(n), np.full(n, 0.3), np.full(n, -0.2),
                        np.full(n, 0.1))
    np.testing.assert_allclose(D[0], 6.0, rtol=1e-12)
    np.testing.assert_allclose(D[5], 6.0, rtol=1e-12)
    np.testing.assert_allclose(D[[1, 2, 3, 4, 6, 7]], 0.0, atol=1e-9)
    Dv = O.tangent_nodes(O.rheo_model(), [O.phase(eta0=3.0)], np.ones((1, n)), np.ones(n),
                         np.ones(n), np.zeros(n), np.full(n, 0.3), np.full(n, -0.2),
                         np.full(n, 0.1), centroid=0)
    np.testing.assert_allclose(Dv[2], 3.0, rtol=1e-12)
    np.testing.assert_allclose(Dv[[0, 1, 3]], 0.0, atol=1e-9)


@pytest.mark.parametrize("centroid", [1, 0])
def test_tangent_power_law_analytic(centroid):
    """eta = B E^(1/n-1) exp(P V/(n R T)): derivatives in closed form."""
    nn, A, V, R, T = 3.0, 2.0, 0.4, 1.0, 1.0
    ph = O.phase(cstv=0, pwlv=1, n_pwl=nn, A_pwl=A, V_pwl=V)
    Exx, Ezz, Exz, P = 0.7, -0.3, 0.25, 0.5
    D = O.tangent_nodes(O.rheo_model(R=R), [ph], np.ones((1, 1)), [1.0], [T], [P], [Exx], [Ezz],
                        [Exz], centroid=centroid)
    E = math.sqrt(0.5 * (Exx ** 2 + Ezz ** 2 + (Exx + Ezz) ** 2) + Exz ** 2)
    B = A ** (-1 / nn) * math.exp(P * V / (nn * R * T))
    eta = B * E ** (1 / nn - 1)
    dEta = B * (1 / nn - 1) * E ** (1 / nn - 2)
    de_xx = dEta * (2 * Exx + Ezz) / (2 * E)
    de_zz = dEta * (Exx + 2 * Ezz) / (2 * E)
    de_g = dEta * Exz / (2 * E)               # w.r.t. gxz = 2 Exz
    de_p = eta * V / (nn * R * T)
    if centroid:
        want = [2 * eta + 2 * de_xx * Exx, 2 * de_zz * Exx, 2 * de_g * Exx, 2 * de_p * Exx,
                2 * de_xx * Ezz, 2 * eta + 2 * de_zz * Ezz, 2 * de_g * Ezz, 2 * de_p * Ezz]
    else:
        want = [2 * de_xx * Exz, 2 * de_zz * Exz, eta + 2 * de_g * Exz, 2 * de_p * Exz]
    np.testing.assert_allclose(D[:, 0], want, rtol=1e-7)
