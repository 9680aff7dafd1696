"""The oracle's Picard assembly and KillerSolver on the reference's SolVi
benchmark (TESTS/SolViBenchmarkTests.cpp): the L2 errors against the
analytic solution stay inside the bounds the reference's test asserts
(tests/golden/reference_known_answers.json), and the velocity error falls
when the grid is refined."""
import json
import os

import numpy as np

from mdgpu import setups as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


def _solvi_errors(O, solvi_bcs, n):
    from oracle.driver import initial_fields, linear_defect_step
    g = O.Grid(n, n, -0.5, 0.5, -0.5, 0.5)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=1)
    ph = ((x * x + z * z) < 0.2 ** 2).astype(np.int32)
    eta = np.array([1.0, 1e3])
    eta_n = O.p2mastah(g, x, z, ph, eta, centroid=1)
    eta_s = O.p2mastah(g, x, z, ph, eta, centroid=0)
    solvi_bcs(g, S)
    num = g.numbering()
    u, v, p = initial_fields(g)
    linear_defect_step(g, num, u, v, p, eta_n, eta_s, np.zeros((n - 1) ** 2), np.zeros(n * n), penalty=1e5,
                       lin_abs_div=1e-10, lin_rel_div=1e-10, lin_abs_mom=1e-10, lin_rel_mom=1e-10,
                       rel_tol_KSP=1e-8)
    XU, ZU = np.meshgrid(g.xg, g.zvx)
    ax, _, _ = S.solvi_analytic(XU.ravel(), ZU.ravel())
    XV, ZV = np.meshgrid(g.xvz, g.zg)
    _, az, _ = S.solvi_analytic(XV.ravel(), ZV.ravel())
    XC, ZC = np.meshgrid(g.xc, g.zc)
    _, _, ap = S.solvi_analytic(XC.ravel(), ZC.ravel())
    L2 = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    return L2(u, ax), L2(v, az), L2(p, ap)


def test_solvi_errors_within_the_reference_bounds(oracle_mod):
    from conftest import solvi_bcs
    ka = json.load(open(GOLDEN))["solvi_51"]
    ex, ez, ep = _solvi_errors(oracle_mod, solvi_bcs, 51)
    print("SolVi 51: L2 Vx %.3e Vz %.3e P %.3e" % (ex, ez, ep))
    assert ex < ka["L2_Vx_max"] and ez < ka["L2_Vz_max"] and ep < ka["L2_P_max"]
    fx, fz, _ = _solvi_errors(oracle_mod, solvi_bcs, 101)
    print("SolVi 101: L2 Vx %.3e Vz %.3e" % (fx, fz))
    assert fx < ex and fz < ez
