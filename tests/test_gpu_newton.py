"""Newton linearisation on the GPU: tangent coefficients against the oracle
(same finite-difference algorithm, FD_Jacobian.c:48-323) and the
matrix-free Jacobian rows against the oracle's assembled Newton Jacobian
(itself pinned in test_oracle_jacobian.py).

Tolerances: tangent 1e-6 relative to each coefficient row's largest value
(the coefficients are differences of viscosities perturbed by 1e-5 or 1e-7
relative, so last-bit differences of device pow/exp are amplified ~1e5-1e7);
J x 1e-13 of the largest entry (summation order)."""
import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S

pytestmark = pytest.mark.gpu


def _setup(O, n, C, seed=4):
    from test_gpu_nonlinear import _problem
    g, st = _problem(O, n, seed=seed)
    phs = S.vep_phases(C=C)
    common = dict(step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4, max_eta=1e4,
                  eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    o_ph = []
    for d in phs:
        d = dict(d)
        b = d.pop("beta")
        o_ph.append(O.OrcPhase(beta_for_K=b, **d))
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    return g, st, ctx, o_ph, [mdgpu.PhaseProp(**d) for d in phs], common


def _perturbed_state(O, g, st, rng, amp=0.05):
    u, v = st["u"], st["v"]
    u += amp * rng.standard_normal(u.size) * (g.BCu_t == -1)
    v += amp * rng.standard_normal(v.size) * (g.BCv_t == -1)
    st["p"] += 0.2 * rng.standard_normal(st["p"].size)
    O.apply_bc(g, u, v)


def _tangents(O, g, st, o_ph, g_ph, common, ctx, average=0):
    from oracle.nonlinear import strain_rate_components
    sr = strain_rate_components(g, st["u"], st["v"])
    P_s = O.c2v(g, st["p"])
    T_s = O.c2v(g, st["T"])
    md = O.OrcRheoModel(**common)
    prm = mdgpu.RheoParams(average=average, **common)
    cin = dict(perc=st["perc_n"], G=st["mu_n"], T=st["T"], P=st["p"], exx=sr["exxd"],
               ezz=sr["ezzd"], exz=sr["exz_n"], sxx0=st["sxxd0"], szz0=st["szzd0"],
               sxz0=st["sxz0_n"])
    vin = dict(perc=st["perc_s"], G=st["mu_s"], T=T_s, P=P_s, exx=sr["exxd_s"],
               ezz=sr["ezzd_s"], exz=sr["exz"], sxx0=st["sxxd0_s"], szz0=st["szzd0_s"],
               sxz0=st["sxz0"])
    Dn_o = O.tangent_nodes(md, o_ph, *[cin[k] for k in ("perc", "G", "T", "P", "exx", "ezz", "exz",
                                                       "sxx0", "szz0", "sxz0")],
                           tag=g.BCp_t, centroid=1, average=average)
    Ds_o = O.tangent_nodes(md, o_ph, *[vin[k] for k in ("perc", "G", "T", "P", "exx", "ezz", "exz",
                                                       "sxx0", "szz0", "sxz0")],
                           tag=g.BCg_t, centroid=0, average=average)
    Dn_g = ctx.tangent_grid(g_ph, prm, 1, **cin)
    Ds_g = ctx.tangent_grid(g_ph, prm, 0, **vin)
    return Dn_o, Ds_o, Dn_g, Ds_g


@pytest.mark.parametrize("centroid", [1, 0])
def test_tangent_all_creep_laws(oracle_mod, centroid):
    """Tangent coefficients of phases with every creep law (power law with
    grain-size and melt factors, diffusion, GBS, Peierls) on random node
    fields against the oracle.  Tolerance: the local iterations stop at a
    1e-13 relative residual, and a derivative by a 1e-5 (centres) or 1e-7
    (vertices) relative perturbation turns that into up to 1e-8 or 1e-6 of
    eta, raised by the stress exponent of the stiffest law (ST + n ~ 10 for
    the Peierls phases here): 1e-6 of each row's largest value on centres,
    2e-5 on vertices."""
    from test_gpu_rheology import LAWS, _ctx, _fields, _phase_pair
    O = oracle_mod
    g, ctx = _ctx(O)
    n = ctx.node_count(centroid)
    pairs = [_phase_pair(O, d) for d in LAWS]
    f, rng = _fields(n, len(LAWS), seed=77 + centroid)
    f["d"] = rng.uniform(0.3, 1.5, n)
    f["phi"] = rng.uniform(0.0, 0.05, n)
    common = dict(step=3, elastic=1, dt=0.5, R=1.0, zeroC=0.6, min_eta=1e-3, max_eta=1e3,
                  eta_tol=1e-12, Tiimin=1e-3, scale_E=1.0, scale_eta=1.0, T_peierls=1.0, d_min=1e-3)
    md = O.OrcRheoModel(**common)
    D_o = O.tangent_nodes(md, [p[0] for p in pairs], *[f[k] for k in ("perc", "G", "T", "P", "exx", "ezz", "exz",
                                                                      "sxx0", "szz0", "sxz0")],
                          d=f["d"], phi=f["phi"], tag=(g.BCp_t if centroid else g.BCg_t), centroid=centroid)
    D_g = ctx.tangent_grid([p[1] for p in pairs], mdgpu.RheoParams(average=0, **common), centroid, **f)
    for r in range(D_o.shape[0]):
        scale = np.abs(D_o[r]).max() + 1e-300
        np.testing.assert_allclose(D_g[r], D_o[r], rtol=0, atol=(1e-6 if centroid else 2e-5) * scale,
                                   err_msg="row %d" % r)


@pytest.mark.parametrize("average", [0, 1, 2])
@pytest.mark.parametrize("C", [1.0, 50.0])
def test_tangent_matches_oracle(oracle_mod, average, C):
    O = oracle_mod
    g, st, ctx, o_ph, g_ph, common = _setup(O, 23, C)
    _perturbed_state(O, g, st, np.random.default_rng(1))
    Dn_o, Ds_o, Dn_g, Ds_g = _tangents(O, g, st, o_ph, g_ph, common, ctx, average)
    for D_o, D_g, eta_row in ((Dn_o, Dn_g, 0), (Ds_o, Ds_g, 2)):
        eta_scale = np.abs(D_o[eta_row]).max()     # D11 ~ 2 eta (centres), D33 ~ eta (vertices)
        for r in range(D_o.shape[0]):
            scale = np.abs(D_o[r]).max() + 1e-300
            # rows 3 and 7 are the pressure derivatives D14, D24, D34: without
            # yielding the viscosity does not depend on P and they are pure
            # finite-difference round-off, a few ulps of eta divided by the
            # perturbation 2e-5 |P| (centres) or 2e-7 |P| (vertices), i.e. up
            # to ~1e-7 of eta here; 1e-6 of the viscosity row bounds them
            noise = 1e-6 * eta_scale if r % 4 == 3 else 0.0
            np.testing.assert_allclose(D_g[r], D_o[r], rtol=1e-6, atol=max(1e-6 * scale, noise),
                                       err_msg="row %d" % r)


@pytest.mark.parametrize("bc", ["pure_shear", "solvi"])
@pytest.mark.parametrize("coeffs", ["random", "tangent"])
def test_jacobian_matches_oracle_assembly(oracle_mod, bc, coeffs):
    """J x on the GPU (matrix-free rows) against the oracle's assembled
    Newton Jacobian (StokesAssemblyDecoupled.c:219-1085,1786-2304, comp = 1)
    times x, with the Picard continuity rows, to 1e-13 of the largest entry:
    every row including the wall-adjacent ones, free-slip and no-slip
    walls, arbitrary signed coefficients or the tangent of a yielding
    visco-elasto-plastic state (C = 1)."""
    import scipy.sparse as sp
    from conftest import solvi_bcs
    O = oracle_mod
    if coeffs == "random":
        n = 27
        g = O.Grid(n, n, 0.0, 1.0, -1.0, 0.0)
        if bc == "pure_shear":
            g.pure_shear(1.0)
        else:
            solvi_bcs(g, S)
        ctx = mdgpu.Context(n, n, 0.0, 1.0, -1.0, 0.0)
        ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
        rng = np.random.default_rng(11)
        Dn = rng.standard_normal((8, (n - 1) ** 2))
        Ds = rng.standard_normal((4, n * n))
    else:
        if bc == "solvi":
            pytest.skip("the yielding set-up is a free-slip box")
        g, st, ctx, o_ph, g_ph, common = _setup(O, 23, C=1.0)
        rng = np.random.default_rng(5)
        _perturbed_state(O, g, st, rng)
        _, _, Dn, Ds = _tangents(O, g, st, o_ph, g_ph, common, ctx)
    num = g.numbering()
    eu, ev, ep, nm, nc, neq = num
    AJ, BJ = O.assemble_jacobian(g, num, Dn, Ds, co=1.0)
    _, _, Cm = O.assemble_picard(g, num, np.ones(Dn.shape[1]), np.ones(Ds.shape[1]))
    K = sp.bmat([[AJ, BJ], [Cm, None]]).tocsr()
    for trial in range(3):
        x = rng.standard_normal(neq)
        y = ctx.jacobian_apply(np.ascontiguousarray(Dn), np.ascontiguousarray(Ds), x)
        ref = K @ x
        err = np.abs(y - ref).max() / np.abs(ref).max()
        print(coeffs, bc, "trial", trial, "max rel. diff.", err)
        assert err <= 1e-13


def test_newton_loop_converges_to_picard_solution(oracle_mod):
    """Picard then Newton (with the forced line search) reaches the same
    non-linear solution as the oracle's Picard iterations driven to a tight
    tolerance, in fewer iterations."""
    from mdgpu.nonlinear import PicardLoop
    from oracle.nonlinear import picard_loop
    O = oracle_mod
    n = 23
    g, st, ctx, o_ph, g_ph, common = _setup(O, n, C=50.0)
    st_gpu = {k: v.copy() for k, v in st.items()}
    tol = (1e-14, 1e-11, 1e-14, 1e-11)
    hist_o, ns_o = picard_loop(g, st, O.OrcRheoModel(**common), o_ph, 60, tol, diag_scaling=1,
                               gz=-1.0, penalty=1e3, max_its_PH=50, rel_tol_KSP=1e-12,
                               max_its_KSP=200, lin_abs_div=1e-14, lin_rel_div=1e-13,
                               lin_abs_mom=1e-14, lin_rel_mom=1e-13)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, g_ph, mdgpu.RheoParams(average=0, **common))
    # inexact Newton: each linear solve to 1e-10 (the loop's own test is on
    # the non-linear residual, 1e-11 relative)
    hist_g, ns_g, stats = loop.run(st_gpu, 30, tol, diag_scaling=1, gz=-1.0, newton=True,
                                   picard2newton_tol=1e-2, max_its_pic=5, rel_tol=1e-10,
                                   max_its=400)
    print("oracle Picard solves", ns_o, "GPU Picard+Newton solves", ns_g, "Newton at",
          st_gpu["newton_steps"], "alphas", st_gpu.get("alphas"))
    print("GPU history", np.array(hist_g)[:, 0])
    assert st_gpu["newton_steps"], "the loop never switched to Newton"
    assert ns_g < ns_o
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(st_gpu["u"], st["u"]) < 1e-8
    assert rel(st_gpu["v"], st["v"]) < 1e-8
    assert rel(st_gpu["p"], st["p"]) < 1e-8


@pytest.mark.parametrize("C", [50.0, 1.0])
def test_killer_iteration_counts_on_newton_system(oracle_mod, C):
    """KillerSolver's iteration on a Newton system (GCR on the Jacobian,
    preconditioned by the Picard factor, Solvers.c:1176-1193): the GPU's
    method 1 (Picard multigrid CG in place of the factor) takes the same
    Powell-Hestenes and GCR iteration counts as the oracle (exact factor)
    and returns the same correction, at a power-law state (C = 50) and at a
    yielding one (C = 1)."""
    O = oracle_mod
    from mdgpu.nonlinear import PicardLoop
    g, st, ctx, o_ph, g_ph, common = _setup(O, 23, C=C)
    rng = np.random.default_rng(9)
    _perturbed_state(O, g, st, rng)
    num = g.numbering()
    eu, ev, ep, nm, nc, neq = num
    loop = PicardLoop(ctx, eu, ev, ep, g_ph, mdgpu.RheoParams(average=0, **common))
    sr, cen, ver = loop.update_nonlinearity(st)
    cin, vin = loop.last_inputs
    Dn = ctx.tangent_grid(g_ph, loop.prm, 1, **cin)
    Ds = ctx.tangent_grid(g_ph, loop.prm, 0, **vin)
    en, es = cen["eta"], ver["eta"]
    rx, rz = ctx.eval_rhs(0.0, -1.0, st["rho_n"], st["rho_s"])
    F, _ = ctx.stokes_residual(cen["sxx"], cen["szz"], ver["sxz"], cen["p_corr"], sr["div"], rx, rz)
    A, B, Cm = O.assemble_picard(g, num, en, es)
    AJ, BJ = O.assemble_jacobian(g, num, Dn, Ds, co=1.0)
    kw = dict(penalty=1e3, max_its_PH=10, rel_tol_KSP=1e-4, max_its_KSP=100, lin_abs_div=1e-9,
              lin_rel_div=1e-5, lin_abs_mom=1e-9, lin_rel_mom=1e-5)
    du, dp, so, rc = O.killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], AJ=AJ, BJ=BJ, **kw)
    sol, o = ctx.newton_solver(Dn, Ds, 2 * en, es, F[:nm], F[nm:], method=1, accept_diverged=True, **kw)
    print("C %g: oracle PH %d KSP %d (status %d); gpu PH %d KSP %d, %d CG its" % (
        C, so.its_PH, so.its_KSP_tot, rc, o.its_PH, o.its_KSP, o.inner_its), flush=True)
    assert (o.its_PH, o.its_KSP) == (so.its_PH, so.its_KSP_tot)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(sol[:nm], du) <= 1e-8 and rel(sol[nm:], dp) <= 1e-8
