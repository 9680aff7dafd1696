"""GPU parity of the grid rheology (NonNewtonianViscosityGrid), the
centre-to-vertex interpolation and the strain-rate node averages against the
oracle.

Tolerance: the two local Newton iterations stop on residual tests
(1e-13 relative for the visco-elastic one, eta_tol for the return map), and
the device pow/exp/log may differ from the host libm in the last bit, so
the iterates can stop one step apart; results agree to 1e-10 relative to
the largest value of each field (the return map stops on an absolute test,
|F| < eta_tol, so a small plastic strain rate carries an absolute error).
"""
import math

import numpy as np
import pytest

import mdgpu

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def _ctx(oracle_mod, nx=37, nz=29):
    box = (0.0, 1.2, -1.0, 0.0)
    g = oracle_mod.Grid(nx, nz, *box)
    g.pure_shear(1.0)
    ctx = mdgpu.Context(nx, nz, *box)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    return g, ctx


PHASES = [
    dict(cstv=1, eta0=1.0, plast=1, C=2.0, fric=math.radians(30)),
    dict(cstv=0, pwlv=1, n_pwl=3.0, A_pwl=0.5, Q_pwl=1.0, V_pwl=0.1, F_pwl=0.8, plast=1,
         C=0.5, fric=0.3, dil=0.1, eta_vp=0.05, n_vp=2.0, beta=20.0),
    dict(cstv=1, eta0=3.0, pwlv=1, n_pwl=2.0, A_pwl=2.0, Slim=0.8, plast=1, C=1.5, fric=0.6),
]


def _phase_pair(oracle_mod, d):
    d = dict(d)
    beta = d.pop("beta", 1e-50)
    full = dict(cstv=0, pwlv=0, plast=0, eta0=1.0, pref_pwl=1.0, F_pwl=1.0, A_pwl=1.0,
                n_pwl=1.0, Q_pwl=0.0, V_pwl=0.0, C=1e50, fric=0.0, dil=0.0, Slim=1e50,
                eta_vp=0.0, n_vp=1.0)
    full.update(d)
    return oracle_mod.OrcPhase(beta_for_K=beta, **full), mdgpu.PhaseProp(beta=beta, **full)


def _fields(n, nph, seed):
    rng = np.random.default_rng(seed)
    perc = rng.dirichlet(np.ones(nph) * 0.7, size=n).T.copy()
    perc[:, rng.random(n) < 0.3] = 0.0
    perc[0][perc.sum(0) == 0] = 1.0
    f = dict(perc=perc, G=rng.uniform(0.5, 2.0, n), T=rng.uniform(0.5, 1.5, n),
             P=rng.uniform(-1.0, 3.0, n), exx=rng.normal(0, 1, n), ezz=rng.normal(0, 1, n),
             exz=rng.normal(0, 1, n), sxx0=rng.normal(0, 0.5, n), szz0=rng.normal(0, 0.5, n),
             sxz0=rng.normal(0, 0.5, n))
    return f, rng


@pytest.mark.parametrize("centroid", [1, 0])
@pytest.mark.parametrize("average", [0, 1, 2])
@pytest.mark.parametrize("step", [0, 3])
def test_viscosity_grid_matches_oracle(oracle_mod, centroid, average, step):
    g, ctx = _ctx(oracle_mod)
    n = ctx.node_count(centroid)
    pairs = [_phase_pair(oracle_mod, d) for d in PHASES]
    f, rng = _fields(n, len(PHASES), seed=10 * step + average + 100 * centroid)
    extra = {}
    if average == 1:   # per-node softening fields and the prefactor multiplier
        extra = dict(fric=rng.uniform(0.1, 0.6, n), C=rng.uniform(0.2, 2.0, n),
                     dil=rng.uniform(0.0, 0.2, n), pref=rng.uniform(0.5, 2.0, n))
    common = dict(step=step, elastic=1, dt=0.5, R=1.0, zeroC=0.6, min_eta=1e-3, max_eta=1e3,
                  eta_tol=1e-12, Tiimin=1e-3, scale_E=1.0, scale_eta=1.0)
    md = oracle_mod.OrcRheoModel(**common)
    prm = mdgpu.RheoParams(average=average, **common)
    ref = oracle_mod.rheology_nodes(md, [p[0] for p in pairs], f["perc"], f["G"], f["T"], f["P"],
                                    f["exx"], f["ezz"], f["exz"], f["sxx0"], f["szz0"], f["sxz0"],
                                    fric=extra.get("fric"), Coh=extra.get("C"),
                                    dil=extra.get("dil"), pref=extra.get("pref"),
                                    tag=(g.BCp_t if centroid else g.BCg_t), centroid=centroid,
                                    average=average)
    assert ref["status"] == 0
    out = ctx.viscosity_grid([p[1] for p in pairs], prm, centroid, **f, **extra)
    if step > 0:
        assert np.any(ref["eII_pl"] > 0) and np.any(ref["eII_pl"] == 0)
    for k in mdgpu.RHEO_OUT:
        np.testing.assert_allclose(out[k], ref[k], rtol=RTOL,
                                   atol=RTOL * np.abs(ref[k]).max() + 1e-300, err_msg=k)


def test_viscosity_grid_device_arrays_and_errors(oracle_mod):
    g, ctx = _ctx(oracle_mod)
    n = ctx.node_count(1)
    pairs = [_phase_pair(oracle_mod, d) for d in PHASES[:2]]
    f, _ = _fields(n, 2, seed=7)
    prm = mdgpu.RheoParams(step=1, elastic=1, average=2, dt=0.5, R=1.0, zeroC=0.0, min_eta=1e-3,
                           max_eta=1e3, eta_tol=1e-12, Tiimin=1e-3, scale_E=1.0, scale_eta=1.0)
    host = ctx.viscosity_grid([p[1] for p in pairs], prm, 1, **f)
    dev = {k: ctx.device(v.size, host=v.ravel()) for k, v in f.items()}
    douts = {k: ctx.device(n, host=np.zeros(n)) for k in ("eta", "sxx")}
    ctx.viscosity_grid([p[1] for p in pairs], prm, 1, outputs=(), **dev, **douts)
    assert np.array_equal(douts["eta"].numpy(), host["eta"])
    assert np.array_equal(douts["sxx"].numpy(), host["sxx"])
    bad = mdgpu.PhaseProp(cstv=1, eta0=1.0, plast=2, beta=1.0)
    with pytest.raises(mdgpu.MdgError):
        ctx.viscosity_grid([bad], prm, 1, **{k: v for k, v in f.items() if k != "perc"},
                           perc=np.ones((1, n)))


# every creep law: power law with grain-size, fugacity and melt factors,
# diffusion (grain-size dependent), grain-boundary sliding and Peierls
# creep, which acts below T_peierls = 1 (half the nodes)
LAWS = [
    dict(cstv=0, pwlv=1, n_pwl=3.0, A_pwl=0.5, Q_pwl=1.0, V_pwl=0.1, F_pwl=0.8, m_pwl=1.0, r_pwl=1.2,
         f_pwl=2.0, a_pwl=20.0, linv=1, F_lin=0.7, A_lin=2.5, n_lin=1.0, m_lin=3.0, Q_lin=1.5, V_lin=0.2,
         r_lin=1.0, f_lin=0.4, a_lin=2.0, gs_ref=0.7, expv=1, G_exp=0.2, S_exp=4.0, Q_exp=6.0, E_exp=3.0,
         n_exp=2.0, plast=1, C=1.2, fric=0.4, dil=0.05, eta_vp=0.02, beta=5.0),
    dict(cstv=1, eta0=2.0, gbsv=1, F_gbs=0.5, A_gbs=1e-3, n_gbs=3.5, m_gbs=2.0, Q_gbs=1.0, linv=1,
         F_lin=1.0, A_lin=1.0, n_lin=1.0, m_lin=0.0, Q_lin=0.5, gs_ref=1.0),
    dict(cstv=0, expv=1, G_exp=0.1, S_exp=8.0, Q_exp=10.0, E_exp=50.0, n_exp=0.0, pwlv=1, n_pwl=3.5,
         A_pwl=1.0, plast=1, C=2.0, fric=0.5),
]


@pytest.mark.parametrize("centroid", [1, 0])
@pytest.mark.parametrize("average", [0, 1, 2])
@pytest.mark.parametrize("step", [0, 3])
def test_viscosity_grid_all_creep_laws(oracle_mod, centroid, average, step):
    g, ctx = _ctx(oracle_mod)
    n = ctx.node_count(centroid)
    pairs = [_phase_pair(oracle_mod, d) for d in LAWS]
    f, rng = _fields(n, len(LAWS), seed=1000 + 10 * step + average + 100 * centroid)
    f["d"] = rng.uniform(0.3, 1.5, n)
    f["phi"] = rng.uniform(0.0, 0.05, n)
    common = dict(step=step, elastic=1, dt=0.5, R=1.0, zeroC=0.6, min_eta=1e-3, max_eta=1e3,
                  eta_tol=1e-12, Tiimin=1e-3, scale_E=1.0, scale_eta=1.0, T_peierls=1.0, d_min=1e-3)
    md = oracle_mod.OrcRheoModel(**common)
    prm = mdgpu.RheoParams(average=average, **common)
    ref = oracle_mod.rheology_nodes(md, [p[0] for p in pairs], f["perc"], f["G"], f["T"], f["P"],
                                    f["exx"], f["ezz"], f["exz"], f["sxx0"], f["szz0"], f["sxz0"],
                                    d=f["d"], phi=f["phi"], tag=(g.BCp_t if centroid else g.BCg_t),
                                    centroid=centroid, average=average)
    assert ref["status"] == 0
    out = ctx.viscosity_grid([p[1] for p in pairs], prm, centroid, **f)
    if step > 0:
        assert np.any(ref["eII_pl"] > 0) and np.any(ref["eII_pl"] == 0)
    for k in mdgpu.RHEO_OUT:
        np.testing.assert_allclose(out[k], ref[k], rtol=RTOL,
                                   atol=RTOL * np.abs(ref[k]).max() + 1e-300, err_msg=k)


def test_viscosity_grid_grain_size_floor(oracle_mod):
    # grain-size dependent diffusion creep below d_min: the reference exits
    # (RheologyDensity.c:506-509), the entry point returns MDG_ERR_ARG
    g, ctx = _ctx(oracle_mod)
    n = ctx.node_count(1)
    f, _ = _fields(n, 1, seed=3)
    f["perc"] = np.ones((1, n))
    f["d"] = np.full(n, 0.5)
    f["d"][n // 2] = 1e-6
    prm = mdgpu.RheoParams(step=1, elastic=1, average=0, dt=0.5, R=1.0, zeroC=0.0, min_eta=1e-3,
                           max_eta=1e3, eta_tol=1e-12, Tiimin=1e-3, scale_E=1.0, scale_eta=1.0, d_min=1e-3)
    ph = _phase_pair(oracle_mod, LAWS[0])[1]
    with pytest.raises(mdgpu.MdgError, match="grain size"):
        ctx.viscosity_grid([ph], prm, 1, **f)
    f["d"][n // 2] = 0.5
    ctx.viscosity_grid([ph], prm, 1, **f)


@pytest.mark.parametrize("periodic", [0, 1])
def test_c2v_matches_oracle(oracle_mod, periodic):
    g, ctx = _ctx(oracle_mod)
    cent = np.random.default_rng(periodic).normal(size=(g.Nx - 1) * (g.Nz - 1))
    ref = oracle_mod.c2v(g, cent, periodic_x=periodic)
    assert np.array_equal(ctx.c2v(cent, periodic_x=periodic), ref)


def test_strain_rate_node_averages(oracle_mod):
    from oracle.driver import initial_fields
    g, ctx = _ctx(oracle_mod)
    rng = np.random.default_rng(5)
    u, v, p = initial_fields(g)
    u += 0.1 * rng.standard_normal(u.size) * (g.BCu_t == -1)
    v += 0.1 * rng.standard_normal(v.size) * (g.BCv_t == -1)
    oracle_mod.apply_bc(g, u, v)
    nv = g.Nx * g.Nz
    sr = oracle_mod.strain_rates(g, u, v)
    exs, ezs, dvs = np.zeros(nv), np.zeros(nv), np.zeros(nv)
    oracle_mod.lib().orc_strain_rates(g.c, 0, u, v, sr["div"], sr["exxd"], sr["ezzd"], sr["exz"],
                                      sr["wxz"].ctypes.data, exs.ctypes.data, ezs.ctypes.data,
                                      dvs.ctypes.data)
    gs = ctx.strain_rates(u, v)
    out = ctx.strain_rates_nodes(gs["exxd"], gs["ezzd"], gs["exz"], gs["div"])
    np.testing.assert_allclose(out["exz_n"], oracle_mod.exz_n(g, sr["exz"]), rtol=0, atol=1e-13)
    for k, r in (("exxd_s", exs), ("ezzd_s", ezs), ("div_s", dvs)):
        np.testing.assert_allclose(out[k], r, rtol=0, atol=1e-13, err_msg=k)


@pytest.mark.parametrize("centroid", [1, 0])
@pytest.mark.parametrize("style", [0, 1, 2])
def test_material_grid_matches_oracle(oracle_mod, centroid, style):
    g, ctx = _ctx(oracle_mod)
    n = ctx.node_count(centroid)
    rng = np.random.default_rng(20 + style + 3 * centroid)
    vals = [dict(G=1.0, bet=2.0, C=1.0, C_end=0.2, phi=0.5, phi_end=0.1, psi=0.2, psi_end=0.0,
                 pls_start=0.1, pls_end=1.1, H_c=-0.3, phi_soft=1, psi_soft=1, coh_soft=1),
            dict(G=3.0, bet=0.5, C=2.0, C_end=1.0, phi=0.6, phi_end=0.3, psi=0.1, psi_end=0.05,
                 pls_start=0.0, pls_end=0.5, H_c=0.2, phi_soft=0, psi_soft=1, coh_soft=1)]
    o_ph = [oracle_mod.OrcPhaseSoft(**v) for v in vals]
    g_ph = [mdgpu.PhaseSoft(G=v["G"], beta=v["bet"], C=v["C"], C_end=v["C_end"], fric=v["phi"],
                            fric_end=v["phi_end"], dil=v["psi"], dil_end=v["psi_end"],
                            pls_start=v["pls_start"], pls_end=v["pls_end"], H_c=v["H_c"],
                            fric_soft=v["phi_soft"], dil_soft=v["psi_soft"],
                            coh_soft=v["coh_soft"]) for v in vals]
    f = rng.uniform(0, 1, n)
    perc = np.stack([f, 1 - f])
    epl = rng.uniform(0, 2, n)
    ref = oracle_mod.material_grid(o_ph, perc, epl, tag=(g.BCp_t if centroid else g.BCg_t),
                                   centroid=centroid, style=style)
    out = ctx.material_grid(g_ph, centroid, perc, epl, style=style)
    for k in ("mu", "bet", "C", "fric", "dil"):
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-14, atol=0, err_msg=k)
