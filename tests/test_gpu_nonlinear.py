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
