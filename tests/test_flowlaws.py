"""Power-law database (ReadDataPowerLaw, FlowLaws.c:46-588) and model scales
(ScaleMe, InputOutput.c:1835-1849), pinned by what the physics fixes:

* the axial-compression correction factor F (t = 1) follows from writing a
  triaxial experiment (strain rate A dsigma^n along the axis, incompressible)
  in second invariants: tau_II = dsigma / sqrt(3), E_II = sqrt(3)/2 * rate,
  tau_II = 2 F A^(-1/n) E_II^(1/n);
* the scaled parameters give, in scaled units, the SI viscosity divided by
  the viscosity scale, for every tabulated law and for any choice of scales
  (a wrong exponent in A's, Q's or V's scaling fails this);
* the user-defined law 1 with n = 1, Q = 0 returns eta0;
* unknown numbers are refused, negative numbers drop F.
The simple-shear factor (t = 2) and the reference-stress law (2), which no
tabulated law uses, are restated without an independent pin."""
import math

import numpy as np
import pytest

from mdgpu import flowlaws as FL

SCALES = [FL.Scale(eta=1e22, L=1e4, V=1e-9, T=600.0), FL.Scale(eta=1e18, L=3e5, V=3e-10, T=1.0)]


@pytest.mark.parametrize("n", [1.0, 2.3, 3.5, 4.7, 18.0])
def test_axial_compression_factor(n):
    A, rate = 3.7e-20, 2e-14
    dsig = (rate / A) ** (1.0 / n)
    tau_II, E_II = dsig / math.sqrt(3.0), math.sqrt(3.0) / 2.0 * rate
    F = tau_II / (2.0 * A ** (-1.0 / n) * E_II ** (1.0 / n))
    assert FL.correction_factor(1, n) == pytest.approx(F, rel=1e-14)
    assert FL.correction_factor(0, n) == 1.0


def _eta_scaled(law, scl, Eii, T, P, R=8.314):
    """What the rheology kernels compute (B Eii^(1/n - 1)) from the scaled
    mdg_phase_prop fields, in scaled units."""
    pp = FL.phase_prop_fields(law)
    Rs, Ts, Ps, Es = R / (scl.J / scl.T), T / scl.T, P / scl.S, Eii / scl.E
    B = pp["pref_pwl"] * pp["F_pwl"] * pp["A_pwl"] ** (-1.0 / pp["n_pwl"]) * \
        math.exp((pp["Q_pwl"] + Ps * pp["V_pwl"]) / Rs / pp["n_pwl"] / Ts) * \
        pp["f_pwl"] ** (-pp["r_pwl"] / pp["n_pwl"])
    return B * Es ** (1.0 / pp["n_pwl"] - 1.0)


def _rate_si(law_si, tau, T, P, d, R=8.314):
    """Laboratory form: strain-rate invariant of a tabulated power / diffusion /
    GBS law at stress invariant tau [Pa], grain size d [m], SI units."""
    t, n, m, r, Q, V, A, f, a = law_si
    return (tau / (2 * FL.correction_factor(t, n))) ** n * A * d ** -m * f ** r * math.exp(-(Q + P * V) / (R * T))


@pytest.mark.parametrize("scl", SCALES)
@pytest.mark.parametrize("kind,number", [("lin", k) for k in sorted(FL.LINEAR_LAWS)] +
                         [("gbs", k) for k in sorted(FL.GBS_LAWS)])
def test_scaled_grain_size_laws(kind, number, scl):
    """Diffusion creep and GBS (ReadDataLinear :591-706, ReadDataGBS
    :711-776): the scaled fields, used as ViscosityConcise uses them
    (RheologyDensity.c:505-515), give the SI strain rate of the published
    law at a stress, temperature, pressure and grain size, for any scales."""
    table = FL.LINEAR_LAWS if kind == "lin" else FL.GBS_LAWS
    law = (FL.linear_law if kind == "lin" else FL.gbs_law)(number, scl)
    pp = (FL.linear_prop_fields if kind == "lin" else FL.gbs_prop_fields)(law)
    g = lambda k: pp[k + "_" + kind]
    tau, T, P, d, R = 3e7, 1300.0, 2e9, 2e-3, 8.314
    Rs, Ts, Ps, ds, taus = R / (scl.J / scl.T), T / scl.T, P / scl.S, d / scl.L, tau / scl.S
    B = g("F") * g("A") ** (-1.0 / g("n")) * math.exp((g("Q") + Ps * g("V")) / Rs / g("n") / Ts) * \
        g("f") ** (-g("r") / g("n")) * ds ** (g("m") / g("n"))
    rate_scaled = (taus / (2.0 * B)) ** g("n")
    assert rate_scaled * scl.E == pytest.approx(_rate_si(table[number][1:], tau, T, P, d), rel=1e-11)
    assert (FL.linear_law if kind == "lin" else FL.gbs_law)(-number, scl)["F"] == 1.0
    with pytest.raises(ValueError):
        (FL.linear_law if kind == "lin" else FL.gbs_law)(7, scl)


@pytest.mark.parametrize("scl", SCALES)
@pytest.mark.parametrize("number", sorted(FL.EXP_LAWS))
def test_scaled_peierls_law(number, scl):
    """ReadDataExponential (:782-842): at sigma = G S the regularised law
    of ViscosityConcise (RheologyDensity.c:516-523), evaluated with the
    scaled fields, gives the SI rate of the exponential law in an axial test
    (sigma = sqrt(3) tau_II, E_II = sqrt(3)/2 rate)."""
    law = FL.exponential_law(number, scl)
    _, _, G, S, Q, V, E, n = FL.EXP_LAWS[number][1:]
    T, R = 900.0, 8.314
    sig = G * S
    rate_si = math.sqrt(3.0) / 2.0 * E * math.exp(-Q / (R * T) * (1.0 - sig / S) ** 2) * sig ** n
    Rs, Ts = R / (scl.J / scl.T), T / scl.T
    ST = law["Q"] / Rs / Ts * 2.0 * law["G"] * (1 - law["G"])
    Arr = math.exp(-law["Q"] / Rs / Ts * (1.0 - law["G"]) ** 2)
    F = (2.0 ** (1.0 - ST - law["n"]) / math.sqrt(3.0) ** (ST + law["n"] + 1.0)) ** (1.0 / (ST + law["n"]))
    B = F * (law["G"] * law["S"]) ** (ST / (ST + law["n"])) / (law["E"] * Arr) ** (1.0 / (ST + law["n"]))
    tau_s = sig / math.sqrt(3.0) / scl.S
    rate_scaled = (tau_s / (2.0 * B)) ** (ST + law["n"])
    assert rate_scaled * scl.E == pytest.approx(rate_si, rel=1e-9)
    assert FL.exponential_law(-number, scl)["t"] == 0.0


@pytest.mark.parametrize("scl", SCALES)
@pytest.mark.parametrize("number", sorted(FL.POWER_LAWS))
def test_scaled_law_gives_si_viscosity(number, scl):
    law = FL.power_law(number, scl)
    si = FL.POWER_LAWS[number][1:]
    Eii, T, P = 1e-14, 1100.0, 1e9
    eta_si = FL.viscosity_si(si, Eii, T, P)
    assert np.isfinite(eta_si) and eta_si > 0
    assert _eta_scaled(law, scl, Eii, T, P) * scl.eta == pytest.approx(eta_si, rel=1e-11)


def test_user_defined_laws_and_errors():
    scl = SCALES[0]
    law = FL.power_law(1, scl, n=1.0, Q=0.0, eta0=1e21 / scl.eta)
    assert law["F"] == 1.0 and law["t"] == 0.0
    assert _eta_scaled(law, scl, 1e-15, 900.0, 0.0) * scl.eta == pytest.approx(1e21, rel=1e-13)
    law3 = FL.power_law(3, scl, n=3.0, Q=2e5, A=1e-20)
    assert law3["A"] == pytest.approx(1e-20 / (scl.S ** -3.0 / scl.t), rel=1e-15)
    assert FL.power_law(-40, scl)["F"] == 1.0
    assert FL.power_law(40, scl)["F"] == pytest.approx(FL.correction_factor(1, 3.5))
    assert FL.power_law(40, scl, force_act_vol_ast=1, act_vol_dis_ast=6e-6)["V"] == \
        pytest.approx(6e-6 / scl.L ** 3)
    assert FL.power_law(34, scl, force_melt_weak=1, melt_weak=20.0)["a"] == 20.0
    for bad in (0, 4, 9, 35, 47):
        with pytest.raises(ValueError):
            FL.power_law(bad, scl)
    # grain-size dependent power laws pass m through to the kernels
    assert FL.phase_prop_fields(dict(FL.power_law(40, scl), m=1.0))["m_pwl"] == 1.0


def test_scale_relations():
    s = FL.Scale(eta=1e22, L=1e4, V=1e-9, T=600.0)
    assert s.t == 1e13 and s.S == pytest.approx(1e9)
    # energy = stress * volume, power = energy / time, density * L^3 = mass
    assert s.J == pytest.approx(s.S * s.L ** 3, rel=1e-15)
    assert s.W == pytest.approx(s.J / s.t, rel=1e-15)
    assert s.rho * s.L ** 3 == pytest.approx(s.m, rel=1e-15)
    assert s.k == pytest.approx(s.W / s.L / s.T, rel=1e-15)
