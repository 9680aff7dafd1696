"""Pins of the oracle's diffusion, grain-boundary-sliding and Peierls creep
and of the power law's grain-size, fugacity and melt factors
(ViscosityConcise, RheologyDensity.c:385-1080; LocalIterationViscoElastic
:152-197) against the published flow-law forms, an independent scalar root
solve and the regularisation's defining properties.  CPU only.

Each law is written here in its laboratory form, strain rate as a function
of stress (FlowLaws.c:591-842 documents the constants), not in the
viscosity form the oracle evaluates:

  power law / diffusion:  e_ii = (tau_ii / (2 F))^n A d^-m f^r exp(a phi) exp(-(Q + P V) / (R T))
  Peierls (Kameyama et al. 1999 regularisation, tangent at sigma = gamma S):
      e_axial = E exp(-Q/(R T) (1 - gamma)^2) (sigma / (gamma S))^s sigma^n,  s = 2 gamma (1 - gamma) Q/(R T),
      axial test: sigma = sqrt(3) tau_ii, e_ii = sqrt(3)/2 e_axial.
"""
import math

import pytest
from scipy.optimize import brentq

import oracle as O

EXX, EZZ, EXZ = 0.7, -0.4, 0.25
EII = math.sqrt(0.5 * (EXX ** 2 + EZZ ** 2 + (EXX + EZZ) ** 2) + EXZ ** 2)


def vc(ph, md, E=1.0, G=1.0, T=1.0, P=0.0, **kw):
    """Oracle at the fixed strain-rate direction scaled to invariant E*EII."""
    return O.viscosity_concise(ph, md, G, T, P, E * EXX, E * EZZ, E * EXZ, **kw)


def rate_power(tau, F, A, n, m=0.0, r=0.0, f=0.0, a=0.0, Q=0.0, V=0.0, d=1.0, phi=0.0, P=0.0, T=1.0, R=1.0):
    out = (tau / (2.0 * F)) ** n * A * math.exp(-(Q + P * V) / (R * T))
    if m:
        out *= d ** -m
    if r:
        out *= f ** r
    if a:
        out *= math.exp(a * phi)
    return out


def peierls_axial(sigma, E, S, gamma, Q, n, T, R=1.0):
    """The exponential (unregularised) Peierls law."""
    return E * math.exp(-Q / (R * T) * (1.0 - sigma / S) ** 2) * sigma ** n


def rate_peierls(tau, E, S, gamma, Q, n, T, R=1.0):
    s = 2.0 * gamma * (1.0 - gamma) * Q / (R * T)
    sig = math.sqrt(3.0) * tau
    return math.sqrt(3.0) / 2.0 * E * math.exp(-Q / (R * T) * (1.0 - gamma) ** 2) * (sig / (gamma * S)) ** s * sig ** n


LIN = dict(linv=1, F_lin=0.7, A_lin=2.5, n_lin=1.0, m_lin=3.0, Q_lin=1.5, V_lin=0.2, r_lin=1.0, f_lin=0.4,
           a_lin=2.0)
PEI = dict(expv=1, G_exp=0.2, S_exp=4.0, Q_exp=6.0, E_exp=3.0, n_exp=2.0)


def test_diffusion_creep_closed_form():
    d, phi, P, T = 0.8, 0.05, 0.3, 1.2
    ph = O.phase(cstv=0, gs_ref=d, **LIN)
    md = O.rheo_model(R=1.0)
    eta, etaVE, _, _, st = vc(ph, md, T=T, P=P, d=d, phi=phi)
    assert st == 0
    tau = 2.0 * etaVE * EII
    e = rate_power(tau, 0.7, 2.5, 1.0, m=3.0, r=1.0, f=0.4, a=2.0, Q=1.5, V=0.2, d=d, phi=phi, P=P, T=T)
    assert e == pytest.approx(EII, rel=1e-13)
    # viscous only, post-processing at the node's own grain size: eta = etaVE
    assert eta == pytest.approx(etaVE, rel=1e-13)


def test_diffusion_dissipation_uses_reference_grain_size():
    # the post-processed viscosity takes d1 = gs_ref (RheologyDensity.c:486, 991)
    d, gs_ref = 0.8, 0.5
    ph = O.phase(cstv=0, gs_ref=gs_ref, **LIN)
    eta, etaVE, *_ = vc(ph, O.rheo_model(R=1.0), d=d)
    assert eta / etaVE == pytest.approx((gs_ref / d) ** 3.0, rel=1e-12)


def test_diffusion_grain_size_floor():
    ph = O.phase(cstv=0, **LIN)
    md = O.rheo_model(R=1.0, d_min=1e-3)
    assert vc(ph, md, d=1e-4)[4] == -2
    assert vc(ph, md, d=1e-2)[4] == 0
    # grain-size independent diffusion creep (m = 0) runs at d = 0
    assert vc(O.phase(cstv=0, **dict(LIN, m_lin=0.0)), md, d=0.0)[4] == 0


@pytest.mark.parametrize("E", [1e-3, 1.0, 30.0])
def test_peierls_tangent_to_exponential_law(E):
    """At sigma = gamma S the regularised law passes through the exponential
    law with its log-slope; the oracle's viscosity reproduces the regularised
    rate at every stress."""
    T = 1.0
    ph = O.phase(cstv=0, **PEI)
    md = O.rheo_model(R=1.0, T_peierls=10.0)
    _, etaVE, _, _, st = vc(ph, md, E=E, T=T)
    assert st == 0
    tau = 2.0 * etaVE * E * EII
    assert rate_peierls(tau, 3.0, 4.0, 0.2, 6.0, 2.0, T) == pytest.approx(E * EII, rel=1e-12)
    # strain rate at which sigma = gamma S, from the exponential law itself
    sig0 = 0.2 * 4.0
    e0 = math.sqrt(3.0) / 2.0 * peierls_axial(sig0, 3.0, 4.0, 0.2, 6.0, 2.0, T)
    _, eta0, *_ = vc(ph, md, E=e0 / EII, T=T)
    assert math.sqrt(3.0) * 2.0 * eta0 * e0 == pytest.approx(sig0, rel=1e-12)
    # log-slope d ln e / d ln sigma there equals the exponential law's
    h = 1e-5
    slope_exact = (math.log(peierls_axial(sig0 * (1 + h), 3.0, 4.0, 0.2, 6.0, 2.0, T)) -
                   math.log(peierls_axial(sig0 * (1 - h), 3.0, 4.0, 0.2, 6.0, 2.0, T))) / \
        (math.log(1 + h) - math.log(1 - h))
    _, ep, *_ = vc(ph, md, E=e0 * (1 + h) / EII, T=T)
    _, em, *_ = vc(ph, md, E=e0 * (1 - h) / EII, T=T)
    tp, tm = 2 * ep * e0 * (1 + h), 2 * em * e0 * (1 - h)
    slope = (math.log(1 + h) - math.log(1 - h)) / (math.log(tp) - math.log(tm))
    assert slope == pytest.approx(slope_exact, rel=1e-6)


def test_peierls_only_below_its_temperature_limit():
    ph = O.phase(cstv=1, eta0=2.0, **PEI)
    md = O.rheo_model(R=1.0, T_peierls=1.5)
    assert vc(ph, md, T=1.6)[:2] == (2.0, 2.0)
    eta, etaVE, *_ = vc(ph, md, T=1.4)
    assert etaVE < 2.0 and eta < 2.0


def test_gbs_changes_only_the_initial_guess():
    # C_gbs is never set in the reference, so sliding adds no strain rate;
    # its viscosity only bounds the starting value of the local iterations
    base = dict(cstv=0, pwlv=1, n_pwl=3.0, A_pwl=2.0, F_pwl=0.6)
    gbs = dict(gbsv=1, F_gbs=0.5, A_gbs=1e-3, n_gbs=3.5, m_gbs=2.0, Q_gbs=1.0)
    md = O.rheo_model(R=1.0, elastic=1, dt=0.4)
    a = vc(O.phase(**base), md, d=0.7)
    b = vc(O.phase(**base, **gbs), md, d=0.7)
    assert b[4] == 0
    assert b[1] == pytest.approx(a[1], rel=1e-12) and b[0] == pytest.approx(a[0], rel=1e-12)


def test_power_law_grain_size_fugacity_melt_factors():
    pw = dict(cstv=0, pwlv=1, n_pwl=3.5, A_pwl=2.0, F_pwl=0.6, Q_pwl=2.0)
    md = O.rheo_model(R=1.0)
    _, e0, *_ = vc(O.phase(**pw), md)
    _, e1, *_ = vc(O.phase(**pw, m_pwl=1.5, r_pwl=1.2, f_pwl=3.0, a_pwl=30.0), md, d=0.5, phi=0.02)
    # e_ii = (tau/2F)^n A d^-m f^r exp(a phi) exp(-Q/RT)
    tau = 2.0 * e1 * EII
    assert rate_power(tau, 0.6, 2.0, 3.5, m=1.5, r=1.2, f=3.0, a=30.0, Q=2.0, d=0.5, phi=0.02) == \
        pytest.approx(EII, rel=1e-12)
    assert e1 / e0 == pytest.approx(0.5 ** (1.5 / 3.5) * 3.0 ** (-1.2 / 3.5) * math.exp(-30.0 * 0.02 / 3.5),
                                    rel=1e-13)


@pytest.mark.parametrize("elastic", [0, 1])
def test_all_laws_match_independent_root(elastic):
    """Constant + power law + diffusion + Peierls (+ Maxwell): the strain
    rates add; tau from brentq on the sum of the laboratory-form rates."""
    G, dt, T, P, d, phi = 1.3, 0.7, 0.9, 0.2, 0.6, 0.01
    ph = O.phase(cstv=1, eta0=5.0, pwlv=1, n_pwl=3.0, A_pwl=1.5, F_pwl=0.6, Q_pwl=1.0, V_pwl=0.5,
                 gs_ref=d, **LIN, **PEI)
    md = O.rheo_model(R=1.0, elastic=elastic, dt=dt, T_peierls=2.0)
    eta, etaVE, _, _, st = vc(ph, md, G=G, T=T, P=P, d=d, phi=phi)
    assert st == 0

    def rates(tau):
        return (tau / (2 * 5.0),
                rate_power(tau, 0.6, 1.5, 3.0, Q=1.0, V=0.5, P=P, T=T),
                rate_power(tau, 0.7, 2.5, 1.0, m=3.0, r=1.0, f=0.4, a=2.0, Q=1.5, V=0.2, d=d, phi=phi, P=P, T=T),
                rate_peierls(tau, 3.0, 4.0, 0.2, 6.0, 2.0, T))

    el = elastic * (lambda tau: tau / (2 * G * dt))(1.0)
    tau = brentq(lambda t: t * el + sum(rates(t)) - EII, 1e-12, 1e3, xtol=1e-15, rtol=1e-15)
    assert 2 * etaVE * EII == pytest.approx(tau, rel=1e-11)
    # dissipative viscosity: the viscous rates only
    assert eta == pytest.approx(tau / (2 * sum(rates(tau))), rel=1e-11)
