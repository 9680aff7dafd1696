"""The oracle's free-slip rows and body force pinned by SolCx.

SolCx (TESTS/SolCxBenchmarkTests.cpp) is a Stokes flow driven by the
density sin(pi z) cos(pi x) under gz = 1 with a 1e6 viscosity jump at
x = 1/2.  Its exact solution (tests/solcx_analytic.py, derived there from
the Stokes equations) is free slip on every wall, so the oracle's closed
free-slip box (walls tag 0, ghost rows tag 13) with the body force of
EvaluateRHS must converge to it: the errors stay inside the bounds the
reference test asserts at 51 x 51 (there with Dirichlet walls) and fall at
the orders it asserts from 41 to 81.  A body force of the wrong sign gives
an error above 1, so the check tells the signs apart.  The analytic solution
itself is checked against its ODE, boundary and interface conditions."""
import json
import os

import numpy as np

from mdgpu import setups as S
from solcx_analytic import SolCx

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")
TIGHT = dict(penalty=1e9, max_its_PH=30, rel_tol_KSP=1e-12, max_its_KSP=500, lin_abs_div=1e-12,
             lin_rel_div=1e-12, lin_abs_mom=1e-12, lin_rel_mom=1e-12)


def solcx_problem(O, n, seed=1):
    """Grid (closed free-slip box), nodal viscosities from 4 x 4 markers per
    cell (arithmetic P2G) and the cell densities of SolCx."""
    g = O.Grid(n, n, 0.0, 1.0, 0.0, 1.0)
    g.pure_shear(0.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    ph = (x >= 0.5).astype(np.int32)
    eta = np.array([1.0, 1e6])
    en = O.p2mastah(g, x, z, ph, eta, centroid=1)
    es = O.p2mastah(g, x, z, ph, eta, centroid=0)
    XC, ZC = np.meshgrid(g.xc, g.zc)
    rho_n = (np.sin(np.pi * ZC) * np.cos(np.pi * XC)).ravel()
    return g, en, es, rho_n


def solcx_errors(g, u, v, p, sc=None):
    """L2 (Vx, Vz on the free nodes) and mean-free L1 (P) errors."""
    sc = sc or SolCx()
    XU, ZU = np.meshgrid(g.xg, g.zvx)
    au, _ = sc.velocity(XU.ravel(), ZU.ravel())
    XV, ZV = np.meshgrid(g.xvz, g.zg)
    _, aw = sc.velocity(XV.ravel(), ZV.ravel())
    XC, ZC = np.meshgrid(g.xc, g.zc)
    ap = sc.pressure(XC.ravel(), ZC.ravel())
    mu, mv = g.BCu_t == -1, g.BCv_t == -1
    L2 = lambda a, b: float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))
    return (L2(u[mu], au[mu]), L2(v[mv], aw[mv]),
            float(np.abs((p - p.mean()) - (ap - ap.mean())).mean()))


def _oracle_solve(O, n, gz=1.0):
    from oracle.driver import initial_fields, linear_defect_step
    g, en, es, rho_n = solcx_problem(O, n)
    num = g.numbering()
    u, v, p = initial_fields(g)
    linear_defect_step(g, num, u, v, p, en, es, rho_n, np.zeros(n * n), gx=0.0, gz=gz, **TIGHT)
    return g, u, v, p


def test_analytic_solution_satisfies_its_equations():
    sc = SolCx()
    h = 2e-3
    for x0 in (0.1, 0.3, 0.45, 0.55, 0.7, 0.9):
        X = sc.X(np.array([x0]))[:, 0]
        Xs = sc.X(x0 + h * np.arange(-2, 3))[0]
        d4 = (Xs[0] - 4 * Xs[1] + 6 * Xs[2] - 4 * Xs[3] + Xs[4]) / h ** 4
        lhs = sc.eta(x0) * (d4 - 2 * np.pi ** 2 * X[2] + np.pi ** 4 * X[0])
        assert abs(lhs + np.pi * np.sin(np.pi * x0)) < 1e-3 * np.pi, x0
        # derivatives returned agree with differences of X
        d1 = (Xs[3] - Xs[1]) / (2 * h)
        assert abs(d1 - X[1]) < 1e-3 * (abs(X[1]) + 1e-6)
    # free slip at the walls: X = X'' = 0 at x = 0 and 1
    Xb = sc.X(np.array([0.0, 1.0 - 1e-15]))
    assert np.all(np.abs(Xb[[0, 2]]) < 1e-12)
    # interface: X, X', eta (pi^2 X + X''), eta (3 pi^2 X' - X''') continuous
    L, R = sc.X(np.array([0.5 - 1e-13]))[:, 0], sc.X(np.array([0.5 + 1e-13]))[:, 0]
    np.testing.assert_allclose(L[:2], R[:2], rtol=1e-6, atol=1e-14)
    np.testing.assert_allclose(1.0 * (np.pi ** 2 * L[0] + L[2]), 1e6 * (np.pi ** 2 * R[0] + R[2]), rtol=1e-6)
    np.testing.assert_allclose(1.0 * (3 * np.pi ** 2 * L[1] - L[3]), 1e6 * (3 * np.pi ** 2 * R[1] - R[3]),
                               rtol=1e-6)


def test_oracle_free_slip_and_body_force_converge_to_solcx(oracle_mod):
    ka = json.load(open(GOLDEN))["solcx"]
    errs, hs = {}, {}
    for n in (41, 51, 81):
        g, u, v, p = _oracle_solve(oracle_mod, n)
        errs[n], hs[n] = solcx_errors(g, u, v, p), g.dx
        print("SolCx %d: L2 Vx %.3e Vz %.3e, L1 P %.3e" % ((n,) + errs[n]))
    ex, ez, ep = errs[51]
    assert ex < ka["L2_Vx_max"] and ez < ka["L2_Vz_max"] and ep < ka["L1_P_max"]
    order = [np.log(errs[41][q] / errs[81][q]) / np.log(hs[41] / hs[81]) for q in range(3)]
    print("orders 41 -> 81: Vx %.2f Vz %.2f P %.2f" % tuple(order))
    assert order[0] >= ka["order_Vx_min"] and order[1] >= ka["order_Vz_min"] and order[2] >= ka["order_P_min"]
    # the body force's sign matters: the opposite gravity gives the opposite flow
    g, u, v, p = _oracle_solve(oracle_mod, 41, gz=-1.0)
    assert min(solcx_errors(g, u, v, p)[:2]) > 1.0
