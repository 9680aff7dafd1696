"""GPU Picard loop on a visco-elasto-plastic shear box against the oracle loop.

Both loops start from the same markers-derived phase fractions, the same
old stresses and the pure-shear boundary velocities, run the non-linear
iterations of Main_DOODZ.c:878-1062 (Picard, alpha = -1 or the line search
of StokesRoutines.c:582-807, diagonal-scaled or plain residual norms) and
must agree on the number of iterations, the chosen step lengths and linear
solves, the residual history (1e-6 relative; the norms of late iterations
are differences of nearly equal stresses), and the final velocity and
pressure to 1e-8 relative L2.
"""
import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S
from mdgpu.nonlinear import PicardLoop

pytestmark = pytest.mark.gpu
TOL = (1e-12, 1e-8, 1e-12, 1e-8)   # abs_mom, rel_mom, abs_div, rel_div


def _problem(O, n, seed):
    from oracle.driver import initial_fields
    box = (0.0, 1.0, -1.0, 0.0)
    g = O.Grid(n + 1, n + 1, *box)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    ph = S.seed_phase(x, z, S.weak_seeds(seed + 2, 6, 2 * g.dx, *box))
    perc_n = O.p2mastah(g, x, z, ph, np.zeros(2), prop=1, centroid=1, nb_phases=2)
    perc_s = O.p2mastah(g, x, z, ph, np.zeros(2), prop=1, centroid=0, nb_phases=2)
    nc, nv = n * n, (n + 1) ** 2
    u, v, p = initial_fields(g)
    rng = np.random.default_rng(seed)
    s0 = 0.2 + 0.05 * rng.standard_normal(nc)
    st = dict(u=u, v=v, p=p, T=np.ones(nc), mu_n=np.ones(nc), mu_s=np.ones(nv), perc_n=perc_n,
              perc_s=perc_s, sxxd0=s0, szzd0=-s0, sxz0_n=np.zeros(nc), sxxd0_s=O.c2v(g, s0),
              szzd0_s=O.c2v(g, -s0), sxz0=np.zeros(nv), rho_n=np.ones(nc), rho_s=np.ones(nv))
    return g, st


@pytest.mark.parametrize("diag_scaling,line_search,C", [(1, False, 1.0), (0, False, 1.0),
                                                        (1, True, 2.0)])
def test_picard_loop_matches_oracle(oracle_mod, diag_scaling, line_search, C):
    O = oracle_mod
    from oracle.nonlinear import picard_loop
    n = 31
    g, st = _problem(O, n, seed=4)
    st_gpu = {k: v.copy() for k, v in st.items()}
    phs = S.vep_phases(C=C)
    common = dict(step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4,
                  eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    o_ph = []
    for d in phs:
        d = dict(d)
        b = d.pop("beta")
        o_ph.append(O.OrcPhase(beta_for_K=b, **d))
    hist, nsolve = picard_loop(g, st, O.OrcRheoModel(**common), o_ph, 10, TOL, average=0,
                               diag_scaling=diag_scaling, gz=-1.0, line_search_on=line_search,
                               penalty=1e3, max_its_PH=50,
                               rel_tol_KSP=1e-12, max_its_KSP=200, lin_abs_div=1e-14,
                               lin_rel_div=1e-12, lin_abs_mom=1e-14, lin_rel_mom=1e-12)
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in phs],
                      mdgpu.RheoParams(average=0, **common))
    ghist, gsolve, stats = loop.run(st_gpu, 10, TOL, diag_scaling=diag_scaling, gz=-1.0,
                                    line_search=line_search, rel_tol=1e-11, max_its=400)
    if line_search:   # LineSearchDecoupled picks the same step lengths
        assert st_gpu["alphas"] == st["alphas"], (st_gpu["alphas"], st["alphas"])
    print("oracle history", np.array(hist)[:, 0], "gpu", np.array(ghist)[:, 0])
    assert gsolve == nsolve and len(ghist) == len(hist)
    np.testing.assert_allclose(np.array(ghist), np.array(hist), rtol=1e-6, atol=1e-15)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(st_gpu["u"], st["u"]) < 1e-8
    assert rel(st_gpu["v"], st["v"]) < 1e-8
    assert rel(st_gpu["p"], st["p"]) < 1e-8
    assert rel(st_gpu["eta_n"], st["eta_n"]) < 1e-8


# every creep law in the loop: the power law of the shear box with
# grain-size and melt factors, plus diffusion creep (grain-size dependent)
# and Peierls creep (active everywhere: T = 1 < T_peierls = 2)
COMMON_LAWS = dict(step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4, eta_tol=1e-13,
                   Tiimin=1e-5, scale_E=1.0, scale_eta=1.0, T_peierls=2.0, d_min=1e-3)


def _law_phases(C=2.0):
    extra = dict(m_pwl=0.5, a_pwl=5.0, linv=1, F_lin=1.0, A_lin=0.5, n_lin=1.0, m_lin=2.0, gs_ref=1.0,
                 expv=1, G_exp=0.2, S_exp=4.0, Q_exp=2.0, E_exp=1.0, n_exp=1.0)
    return [dict(p, **extra) for p in S.vep_phases(C=C)]


def _law_state(O, g, st, seed=11):
    """Grain size (0.5-1.5) and melt fraction (0-0.02) on the centres, carried
    to the vertices as the reference carries centre fields."""
    rng = np.random.default_rng(seed)
    nc = (g.Nx - 1) * (g.Nz - 1)
    st["d0_n"] = rng.uniform(0.5, 1.5, nc)
    st["phi0_n"] = rng.uniform(0.0, 0.02, nc)
    st["d0_s"], st["phi0_s"] = O.c2v(g, st["d0_n"]), O.c2v(g, st["phi0_n"])


def test_picard_loop_all_creep_laws_matches_oracle(oracle_mod):
    """The loop with the line search on phases carrying every implemented
    creep law and per-node grain size and melt fraction (device-resident
    fields): the same steps, solves and residual history as the oracle's
    loop, u, v, p to 1e-8."""
    O = oracle_mod
    from oracle.nonlinear import picard_loop
    n = 31
    g, st = _problem(O, n, seed=4)
    _law_state(O, g, st)
    st_gpu = {k: v.copy() for k, v in st.items()}
    phs = _law_phases()
    o_ph = []
    for d in phs:
        d = dict(d)
        b = d.pop("beta")
        o_ph.append(O.OrcPhase(beta_for_K=b, **d))
    hist, nsolve = picard_loop(g, st, O.OrcRheoModel(**COMMON_LAWS), o_ph, 10, TOL, diag_scaling=1, gz=-1.0,
                               line_search_on=True, penalty=1e3, max_its_PH=50, rel_tol_KSP=1e-12, max_its_KSP=200,
                               lin_abs_div=1e-14, lin_rel_div=1e-12, lin_abs_mom=1e-14, lin_rel_mom=1e-12)
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in phs],
                      mdgpu.RheoParams(average=0, **COMMON_LAWS), device=True)
    ghist, gsolve, _ = loop.run(st_gpu, 10, TOL, diag_scaling=1, gz=-1.0, line_search=True, rel_tol=1e-11,
                                max_its=400)
    print("alphas", st["alphas"], "history", np.array(hist)[:, 0] / hist[0][0], flush=True)
    assert st_gpu["alphas"] == st["alphas"] and gsolve == nsolve and len(ghist) == len(hist)
    np.testing.assert_allclose(np.array(ghist), np.array(hist), rtol=1e-6, atol=1e-15)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    for k in ("u", "v", "p", "eta_n"):
        assert rel(st_gpu[k], st[k]) < 1e-8, k


def test_predictor_line_search_matches_oracle(oracle_mod):
    """pc_line_search (StokesRoutines.c:636-668): the longest step is tried
    first; the GPU loop takes the same steps as the oracle's, including the
    iterations where the predictor wins over the scan's choice."""
    O = oracle_mod
    from oracle.nonlinear import picard_loop
    n = 31
    g, st = _problem(O, n, seed=4)
    st_gpu = {k: v.copy() for k, v in st.items()}
    phs = S.vep_phases(C=2.0)
    common = dict(step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4,
                  eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    o_ph = []
    for d in phs:
        d = dict(d)
        b = d.pop("beta")
        o_ph.append(O.OrcPhase(beta_for_K=b, **d))
    hist, nsolve = picard_loop(g, st, O.OrcRheoModel(**common), o_ph, 10, TOL, diag_scaling=1, gz=-1.0,
                               line_search_on=True, pc_line_search=1, penalty=1e3, max_its_PH=50,
                               rel_tol_KSP=1e-12, max_its_KSP=200, lin_abs_div=1e-14, lin_rel_div=1e-12,
                               lin_abs_mom=1e-14, lin_rel_mom=1e-12)
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in phs], mdgpu.RheoParams(average=0, **common),
                      device=True)
    ghist, gsolve, _ = loop.run(st_gpu, 10, TOL, diag_scaling=1, gz=-1.0, line_search=True, pc_line_search=1,
                                rel_tol=1e-11, max_its=400)
    print("alphas", st["alphas"], flush=True)
    assert -2.5 in st["alphas"]
    assert st_gpu["alphas"] == st["alphas"] and gsolve == nsolve
    np.testing.assert_allclose(np.array(ghist), np.array(hist), rtol=1e-6, atol=1e-15)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(st_gpu["u"], st["u"]) < 1e-8 and rel(st_gpu["p"], st["p"]) < 1e-8


def test_safe_mode_restart_on_the_device(oracle_mod):
    """Safe mode (Main_DOODZ.c:1021-1052) on the device-resident loop: a
    stagnation (forced on the second line search from the host) divides dt by
    safe_dt_div and restarts from the call's u, v, p; the run then equals a
    fresh run at the reduced step (same steps, fields to 1e-12)."""
    n = 31
    g, st0 = _problem(oracle_mod, n, seed=4)
    phs = S.vep_phases(C=2.0)
    common = dict(step=1, elastic=1, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4, eta_tol=1e-13, Tiimin=1e-5,
                  scale_E=1.0, scale_eta=1.0)
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    eu, ev, ep = g.numbering()[:3]
    runs = []
    for dt, stagnate in ((1.0, True), (0.25, False)):
        st = {k: v.copy() for k, v in st0.items()}
        loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in phs],
                          mdgpu.RheoParams(average=0, dt=dt, **common), device=True)
        if stagnate:
            real, calls = loop.line_search, []

            def scripted(*a, **k):
                calls.append(1)
                return (0.0, True) if len(calls) == 2 else real(*a, **k)
            loop.line_search = scripted
        out = loop.run(st, 8, TOL, diag_scaling=1, gz=-1.0, line_search=True, safe_mode=1, safe_dt_div=4.0,
                       rel_tol=1e-11, max_its=400)
        runs.append((st, loop, out))
    (s1, l1, (h1, n1, _)), (s2, l2, (h2, n2, _)) = runs
    assert s1["safe_restarts"] == [1] and l1.prm.dt == 0.25 and s2["safe_restarts"] == []
    assert n1 == n2 + 2 and s1["alphas"][2:] == s2["alphas"]
    np.testing.assert_allclose(np.array(h1[2:]), np.array(h2), rtol=1e-10, atol=1e-15)
    for k in ("u", "v", "p"):
        assert np.linalg.norm(s1[k] - s2[k]) <= 1e-12 * np.linalg.norm(s2[k]), k


@pytest.mark.parametrize("newton", [False, True])
def test_device_resident_loop_matches_host_loop(oracle_mod, newton):
    """The loop with every field in device memory (one upload per run, the
    defect step and line-search trials by mdg_extract_solution) against the
    host-staged loop, on the 31^2 visco-elasto-plastic box with the line
    search (Picard) and on the power-law ShearTemplate case with Newton steps
    from the first iteration: the same iterations, step lengths and solver
    iteration counts, the residual history and the fields to 1e-10 (the
    device update fuses u + alpha dx into one rounding where numpy rounds
    twice)."""
    if newton:
        runs = []
        for device in (False, True):
            st, loop = _shear_template(oracle_mod, False)
            loop.device = device
            out = loop.run(st, 10, (1e-9, 1e-9, 1e-9, 1e-9), diag_scaling=0, newton=True, picard2newton=False,
                           rel_tol=1e-11, max_its=400)
            runs.append((st, out))
    else:
        n = 31
        g, st0 = _problem(oracle_mod, n, seed=4)
        phs = S.vep_phases(C=2.0)
        common = dict(step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4,
                      eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
        ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
        ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
        eu, ev, ep = g.numbering()[:3]
        runs = []
        for device in (False, True):
            st = {k: v.copy() for k, v in st0.items()}
            loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in phs],
                              mdgpu.RheoParams(average=0, **common), device=device)
            out = loop.run(st, 10, TOL, diag_scaling=1, gz=-1.0, line_search=True, rel_tol=1e-11, max_its=400)
            runs.append((st, out))
    (sh, (hh, nh, oh)), (sd, (hd, nd, od)) = runs
    print("host ms per iteration", [round(t, 1) for t in sh["iter_ms"]], "device", [round(t, 1) for t in sd["iter_ms"]],
          "upload %.1f ms" % sd["upload_ms"], flush=True)
    assert nd == nh and len(hd) == len(hh)
    assert sd.get("alphas") == sh.get("alphas") and sd["newton_steps"] == sh["newton_steps"]
    assert [o.its for o in od] == [o.its for o in oh]
    np.testing.assert_allclose(np.array(hd), np.array(hh), rtol=1e-10, atol=1e-15)
    for k in ("u", "v", "p", "eta_n", "eta_s"):
        assert np.linalg.norm(sd[k] - sh[k]) <= 1e-10 * np.linalg.norm(sh[k]), k
    if newton:
        assert sd["newton_steps"]


def _shear_template(O, linear):
    """The ShearTemplate cases of the reference's NewtonIterationConvergence
    test (TESTS/NewtonIterationConvergence.cpp:97-142, parameters
    TESTS/ShearTemplate/LinearPureshearIsotropic.txt and
    NonLinearPureshearIsotropic.txt): 21 x 21 vertices on [-0.5, 0.5]^2,
    4 x 4 markers per cell, pure shear at a background rate of 1, a circular
    inclusion of radius 0.05 (eta = 10, constant) in a matrix of eta = 1
    (linear case) or a 'homemade' power law eta = Eii^(1/3 - 1) (pwlv = 1,
    n = 3, eta0 = 1: A = eta0^-n, FlowLaws.c:58-70); no elasticity, no
    plasticity (C = 1e90), min_eta 1e-3, max_eta 1e6; initial velocity the
    background pure shear (GridRoutines.c:355-385)."""
    box = (-0.5, 0.5, -0.5, 0.5)
    n = 20
    g = O.Grid(n + 1, n + 1, *box)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=1)
    ph = ((x * x + z * z) < 0.05 ** 2).astype(np.int32)
    perc_n = O.p2mastah(g, x, z, ph, np.zeros(2), prop=1, centroid=1, nb_phases=2)
    perc_s = O.p2mastah(g, x, z, ph, np.zeros(2), prop=1, centroid=0, nb_phases=2)
    Nx = n + 1
    u = np.zeros(Nx * (Nx + 1))
    v = np.zeros((Nx + 1) * Nx)
    for l in range(Nx + 1):
        u[l * Nx:(l + 1) * Nx] = -np.asarray(g.xg)
    for l in range(Nx):
        v[l * (Nx + 1):(l + 1) * (Nx + 1)] = g.zg[l]
    for arr, t, val in ((u, g.BCu_t, g.BCu_v), (v, g.BCv_t, g.BCv_v)):
        m = t == 0
        arr[m] = val[m]
    O.apply_bc(g, u, v)
    nc, nv = n * n, Nx * Nx
    z0, z0s = np.zeros(nc), np.zeros(nv)
    st = dict(u=u, v=v, p=np.zeros(nc), T=np.full(nc, 274.15), mu_n=np.ones(nc), mu_s=np.ones(nv),
              perc_n=perc_n, perc_s=perc_s, sxxd0=z0.copy(), szzd0=z0.copy(), sxz0_n=z0.copy(),
              sxxd0_s=z0s.copy(), szzd0_s=z0s.copy(), sxz0=z0s.copy(), rho_n=np.ones(nc),
              rho_s=np.ones(nv))
    common = dict(cstv=0, pwlv=0, plast=0, eta0=1.0, pref_pwl=1.0, F_pwl=1.0, A_pwl=1.0,
                  n_pwl=3.0, Q_pwl=0.0, V_pwl=0.0, C=1e90, fric=np.radians(30.0), dil=0.0,
                  Slim=1e100, eta_vp=0.0, n_vp=1.0, beta=1e-11)
    matrix = dict(common, cstv=1) if linear else dict(common, pwlv=1)
    inclusion = dict(common, cstv=1, eta0=10.0)
    ctx = mdgpu.Context(Nx, Nx, *box)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    prm = mdgpu.RheoParams(step=1, elastic=0, average=0, dt=0.5, R=8.314, zeroC=273.15,
                           min_eta=1e-3, max_eta=1e6, eta_tol=1e-11, Tiimin=0.0, scale_E=1.0,
                           scale_eta=1.0)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**matrix), mdgpu.PhaseProp(**inclusion)], prm)
    return st, loop


@pytest.mark.parametrize("linear", [True, False])
def test_shear_template_newton_iteration_counts(oracle_mod, linear):
    """Known answers of the reference's ShearTemplate tests: Newton with the
    line search (Newton = 1, Picard2Newton = 0, nit_max = 10, non-linear
    tolerances 1e-9, diag_scaling = 0) takes exactly 1 step for the linear
    case and more than 1 and fewer than 10 for the power-law case
    (NewtonIterationConvergence.cpp:104 and :141-142)."""
    st, loop = _shear_template(oracle_mod, linear)
    tol = (1e-9, 1e-9, 1e-9, 1e-9)
    hist, nsolve, stats = loop.run(st, 10, tol, diag_scaling=0, newton=True, picard2newton=False,
                                   rel_tol=1e-11, max_its=400)
    h = np.array(hist)
    print("linear" if linear else "power law", "steps", nsolve, "history", h[:, 0],
          "alphas", st.get("alphas"))
    last, first = h[-1], h[0]
    converged = all(last[i] < tol[0] or last[i] / first[i] < tol[1] for i in range(3))
    assert converged
    import json
    import os
    ka = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                     "reference_known_answers.json")))["shear_template_newton"]
    if linear:
        assert nsolve == ka["linear_newton_steps"]
    else:
        assert ka["power_law_newton_steps_above"] < nsolve < ka["power_law_newton_steps_below"]


def test_unconverged_newton_solves_carry_on(oracle_mod):
    """Newton on the heavily yielding shear-band box (127^2, C = 1) with
    KillerSolver's own iteration (method 1: GCR on the Jacobian preconditioned
    by the Picard operator, as Solvers.c:1176-1193,1298 precondition it with
    the Picard factor).  The Picard solves before the switch converge in a few
    Powell-Hestenes iterations; on the plastic Jacobian GCR stops at its
    iteration caps far from the momentum tolerance, which KillerSolver accepts
    (only the continuity bound aborts, Solvers.c:1333-1336), and the loop
    carries on with the steps the line search weighs.  The Powell-Hestenes
    cap is the reference's default (max_its_PH = 10, InputOutput.c:1184):
    with 3 the continuity misses 10 lin_abs_div on these solves, which the
    reference would abort on as the GPU solver does."""
    from shearbox import build
    ctx, loop, st, info = build(oracle_mod, 127, C=1.0)
    hist, nsolve, stats = loop.run(st, 7, (1e-12, 1e-8, 1e-12, 1e-8), diag_scaling=1, line_search=True,
                                   newton=True, method=1, max_its_PH=10, max_its_KSP=20)
    h = np.array(hist)
    print("newton steps", st["newton_steps"], "alphas", st["alphas"], "history", (h / h[0])[:, :3].tolist(),
          "rel", [o.rel_res for o in stats], flush=True)
    assert st["newton_steps"]
    for nit, o in enumerate(stats):
        assert o.status == 0
        if nit in st["newton_steps"]:
            assert o.rel_res > 1e-3                      # the plastic Jacobian is not solved
        else:
            assert o.rel_res < 1e-8
        assert abs(o.max_div) <= 10 * 1e-9 or o.max_div / o.max_div0 <= 1e-5
    assert np.isfinite(h).all() and h[-1, 0] < h[0, 0]
