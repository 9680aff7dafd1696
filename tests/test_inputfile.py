"""Model-file reader (ReadInputFile and its ReadDou2 / ReadInt2 / ReadChar /
ReadMatProps lookups, InputOutput.c:1077-1560, 1873-2350) on three of the
reference's own test inputs, kept as fixtures under tests/golden/inputs
(TESTS/Thermal/GaussianDiffusionL2.txt, TESTS/RheologyCreep/PowerLawCreep.txt,
TESTS/SolViBenchmark/SolViRes101.txt), and on small texts for the lookup
rules.  The olivine law of PowerLawCreep is checked end to end: reader,
flow-law database and the oracle's ViscosityConcise give, back in Pa s, the
viscosity of the published law written in SI units."""
import math
import os

import pytest

from mdgpu import flowlaws as FL
from mdgpu import inputfile as IF

HERE = os.path.join(os.path.dirname(__file__), "golden", "inputs")


def _model(name):
    return IF.Model.read(os.path.join(HERE, name))


def test_gaussian_diffusion_file():
    m = _model("GaussianDiffusionL2.txt")
    s = m.scaling
    assert (s.eta, s.L, s.V, s.T) == (1e22, 1e4, 1e-10, 500.0)
    assert m["Nx"] == 51 and m["Nz"] == 51 and m["Nt"] == 5
    assert m["xmin"] == pytest.approx(-5e4 / 1e4) and m["zmax"] == pytest.approx(5.0)
    assert m["dt"] == pytest.approx(1e12 / s.t)
    assert m["thermal"] == 1 and m["advection"] == 0
    # Newton = 1 forces the line search and KillerSolver (InputOutput.c:1334-1341)
    assert m["Newton"] == 1 and m["line_search"] == 1 and m["lin_solver"] == 2
    assert m["penalty"] == 1e5 and m["lin_rel_div"] == 1e-9 and m["rel_tol_KSP"] == 1e-4
    assert m["user0"] == 500.0 and m["user3"] == 10000.0
    assert m.writer_subfolder == "GaussianDiffusionL2"
    assert len(m.phases) == 1
    # unset keys keep the reference defaults
    assert m["max_its_PH"] == 10 and m["thermal_solver"] == 0 and m["rel_tol_thermal"] == 1e-8
    assert m["max_its_thermal"] == 1000


def test_solvi_file():
    m = _model("SolViRes101.txt")
    assert m["Nx"] == 101 and m["penalty"] == 1e5 and m["lin_abs_div"] == 1e-8
    assert m["user1"] == 0.2
    props = m.phase_props()
    assert [p.cstv for p in props] == [1, 1]
    assert props[0].eta0 * m.scaling.eta == pytest.approx(1.0)
    assert props[1].eta0 * m.scaling.eta == pytest.approx(1000.0)
    opts = m.solver_opts()
    assert opts.method == 1 and opts.penalty == 1e5 and opts.lin_rel_div == 1e-8


def test_power_law_creep_olivine(oracle_mod):
    O = oracle_mod
    m = _model("PowerLawCreep.txt")
    s = m.scaling
    ph = m.phases[0]
    assert ph.pwlv == 40 and ph.pwl["n"] == 3.5
    assert ph.pwl["V"] == pytest.approx(11.0e-6 / s.L ** 3)
    pp = m.phase_props()[0]
    assert pp.pwlv == 1 and pp.cstv == 0 and pp.F_pwl == pytest.approx(FL.correction_factor(1, 3.5))
    md = m.rheo_params()
    orc_md = O.rheo_model(step=md.step, elastic=md.elastic, dt=md.dt, R=md.R, zeroC=md.zeroC,
                          min_eta=md.min_eta, max_eta=md.max_eta, eta_tol=md.eta_tol, Tiimin=md.Tiimin,
                          scale_E=md.scale_E, scale_eta=md.scale_eta)
    orc_ph = O.phase(**{k: getattr(pp, k) for k, _ in pp._fields_ if k != "beta"}, beta_for_K=pp.beta)
    T_K, E = 600.0 + 273.15, 1e-14
    args = (ph.G, T_K / s.T, 0.0, E / s.E, -E / s.E, 0.0)
    # the file's cut-offs (min_eta = 1e-6, max_eta = 1e6) are divided by the
    # viscosity scale like any other, i.e. read as Pa s: the olivine matrix
    # is clipped to 1e6 Pa s, in the reference run as here
    eta, *_ = O.viscosity_concise(orc_ph, orc_md, *args)
    assert eta * s.eta == pytest.approx(1e6, rel=1e-14)
    orc_md.max_eta = 1e50
    eta, *_ = O.viscosity_concise(orc_ph, orc_md, *args)
    eta_si = FL.viscosity_si(FL.POWER_LAWS[40][1:], E, T_K, 0.0, R=IF.RG)
    assert eta * s.eta == pytest.approx(eta_si, rel=1e-10)
    assert 1e20 < eta_si < 1e24


def test_lookup_rules():
    text = "\n".join([
        "/**** top ****/",
        "Nx=99",                       # no space: not a key line
        "  Nz = 7",                    # leading space: not a key line
        "dt  = 2.5e3  / a comment",
        "penalty = 1e3xyz",
        "Nb_phases = 2",
        "ID = 0",
        "rho = 3000",
        "eta0 = 1e21",
        "ID = 1",
        "rho = 2500",
        "Nz = 42",                     # first match anywhere in the file
        "writer_subfolder = out_dir extra",
    ]) + "\n"
    f = IF.InputFile(text)
    assert f.int("Nx", 10) == 10
    assert f.int("Nz", 10) == 42
    assert f.dou("dt", 0.0) == 2.5e3
    assert f.dou("penalty", 0.0) == 1e3
    assert f.mat("rho", 0, 1.0) == 3000.0 and f.mat("rho", 1, 1.0) == 2500.0
    assert f.mat("eta0", 1, -1.0) == -1.0          # stops at the next ID / end of file
    assert f.mat("eta0", 0, -1.0) == 1e21
    assert f.char("writer_subfolder", "./") == "out_dir"
    assert f.char("missing", "./") == "./"
    with pytest.raises(ValueError):
        f.mat("rho", 5, 1.0)
    assert IF.c_atof("  -3.5e-2abc") == -3.5e-2 and IF.c_atof("abc") == 0.0
    assert IF.c_atoi(" 000005") == 5 and IF.c_atoi("1e3") == 1


# ---- the reference's creep tests (TESTS/RheologyCreep, RheologyCreepTests.cpp:104-213):
# one phase of each file at the file's temperature (user0, degC), background
# strain rate (pure shear) and grain size (SetGrainSize returns gs_ref,
# RheologyCreepTests.cpp:49-51), against the published law in SI units

def _creep_case(O, name, phase=0, lift_cap=True):
    m = _model(name)
    s = m.scaling
    pp = m.phase_props()[phase]
    md = m.rheo_params()
    orc_md = O.rheo_model(**{k: getattr(md, k) for k, _ in md._fields_ if k != "average"})
    if lift_cap:
        orc_md.max_eta, orc_md.min_eta = 1e50, 1e-50
    orc_ph = O.phase(**{k: getattr(pp, k) for k, _ in pp._fields_ if k != "beta"}, beta_for_K=pp.beta)
    T_K, E = m["user0"] + IF.ZERO_C, m["bkg_strain_rate"] * s.E    # SI
    out = O.viscosity_concise(orc_ph, orc_md, m.phases[phase].G, T_K / s.T, 0.0, E / s.E, -E / s.E, 0.0,
                              d=m.phases[phase].gs_ref)
    return m, out, T_K, E


def _rate_si(law_si, tau, T, d):
    t, n, m, r, Q, V, A, f, a = law_si
    return (tau / (2 * FL.correction_factor(t, n))) ** n * A * d ** -m * math.exp(-Q / (IF.RG * T))


def test_reference_creep_inputs_read():
    m = _model("DiffusionCreepLinear.txt")
    ph, pp = m.phases[0], m.phase_props()[0]
    assert (ph.linv, ph.gbsv, ph.expv, pp.linv, pp.pwlv, pp.cstv) == (40, 0, 0, 1, 0, 0)
    assert ph.gs_ref * m.scaling.L == pytest.approx(5e-3)
    assert pp.n_lin == 1.0 and pp.m_lin == 3.0 and pp.V_lin == pytest.approx(4e-6 / m.scaling.L ** 3)
    assert pp.F_lin == pytest.approx(1.0 / 3.0)
    md = m.rheo_params()
    assert md.T_peierls * m.scaling.T == pytest.approx(1473.15) and md.d_min * m.scaling.L == pytest.approx(1e-13)
    pe = _model("PeierlsCreep.txt").phase_props()[0]
    assert pe.expv == 1 and pe.G_exp == 0.1 and pe.n_exp == 0.0
    assert _model("GBSCreep.txt").phase_props()[0].gbsv == 1
    cp = _model("CompositePwlLin.txt").phase_props()[0]
    assert cp.pwlv == 1 and cp.linv == 1


def test_diffusion_creep_input(oracle_mod):
    m, (eta, etaVE, _, _, st), T_K, E = _creep_case(oracle_mod, "DiffusionCreepLinear.txt")
    assert st == 0
    law = FL.LINEAR_LAWS[40][1:]
    tau = 2 * etaVE * m.scaling.eta * E
    assert _rate_si(law, tau, T_K, 5e-3) == pytest.approx(E, rel=1e-10)
    assert eta == pytest.approx(etaVE, rel=1e-12)      # d = gs_ref
    # with the file's cut-offs (read in Pa s like the power-law case) the run is capped
    _, capped, *_ = _creep_case(oracle_mod, "DiffusionCreepLinear.txt", lift_cap=False)
    assert capped[1] * m.scaling.eta == pytest.approx(1e6, rel=1e-14)


def test_diffusion_creep_grain_size_input(oracle_mod):
    # two olivine phases with gs_ref 5 mm and 1 mm: n = 1, m = 3
    _, a, *_ = _creep_case(oracle_mod, "DiffusionCreepGrainSize.txt", 0)
    _, b, *_ = _creep_case(oracle_mod, "DiffusionCreepGrainSize.txt", 1)
    assert a[1] / b[1] == pytest.approx(5.0 ** 3, rel=1e-12)


def test_peierls_creep_input(oracle_mod):
    m, (eta, etaVE, _, _, st), T_K, E = _creep_case(oracle_mod, "PeierlsCreep.txt")
    assert st == 0 and T_K < 1473.15
    _, _, G, S, Q, _, E0, n = FL.EXP_LAWS[40][1:]
    tau = 2 * etaVE * m.scaling.eta * E
    s_ = 2 * G * (1 - G) * Q / (IF.RG * T_K)
    sig = math.sqrt(3.0) * tau
    rate = math.sqrt(3.0) / 2 * E0 * math.exp(-Q / (IF.RG * T_K) * (1 - G) ** 2) * (sig / (G * S)) ** s_ * sig ** n
    assert rate == pytest.approx(E, rel=1e-9)
    assert eta == pytest.approx(etaVE, rel=1e-12)


def test_gbs_creep_input(oracle_mod):
    # grain-boundary sliding adds no strain rate in ViscosityConcise (C_gbs = 0):
    # with elasticity on, the phase is a Maxwell spring G dt, and its
    # dissipative viscosity is 1/0, capped at max_eta
    m, (eta, etaVE, _, _, st), _, _ = _creep_case(oracle_mod, "GBSCreep.txt")
    assert m["elastic"] == 1 and st == 0
    assert etaVE == pytest.approx(m.phases[0].G * m["dt"], rel=1e-12)
    assert eta == 1e50


def test_composite_creep_input(oracle_mod):
    from scipy.optimize import brentq
    m, (eta, etaVE, _, _, st), T_K, E = _creep_case(oracle_mod, "CompositePwlLin.txt")
    assert st == 0
    pw, li = FL.POWER_LAWS[40][1:], FL.LINEAR_LAWS[40][1:]
    tau = brentq(lambda t: _rate_si(pw, t, T_K, 1.0) + _rate_si(li, t, T_K, 5e-3) - E, 1.0, 1e12, rtol=1e-15)
    assert 2 * etaVE * m.scaling.eta * E == pytest.approx(tau, rel=1e-10)
    # diffusion carries ~1 % of the strain rate at 5 mm: dropping it would move tau by ~3e-3
    assert 1e-3 < _rate_si(li, tau, T_K, 5e-3) / E < 0.5


def test_loop_options():
    m = _model("GaussianDiffusionL2.txt")
    nit_max, tol, kw = m.loop_options()
    assert kw["newton"] and kw["line_search"]            # Newton = 1 forces the line search
    assert (kw["safe_mode"], kw["safe_dt_div"], kw["max_num_stag"], kw["pc_line_search"]) == (0, 5.0, 3, 0)
    assert nit_max == m["nit_max"] and tol == (m["nonlin_abs_mom"], m["nonlin_rel_mom"], m["nonlin_abs_div"],
                                               m["nonlin_rel_div"])
    f = IF.InputFile("safe_mode = 1\nsafe_dt_div = 2.5\nmax_num_stag = 4\npc_line_search = 1\nline_search_min = 0.1\n"
                     "Nb_phases = 0\n")
    _, _, kw = IF.Model(f).loop_options()
    assert (kw["safe_mode"], kw["safe_dt_div"], kw["max_num_stag"], kw["pc_line_search"], kw["line_search_min"]) == \
        (1, 2.5, 4, 1, 0.1)
