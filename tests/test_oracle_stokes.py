"""The oracle's strain rates, Picard operator and KillerSolver pinned by what
the discretisation fixes: a linear velocity field has constant strain rates
(the compressible form removes a third of the divergence from the normal
rates, the incompressible one keeps them; the shear rate is the mean of the
cross derivatives) at every interior node; the velocity block of the Picard
operator is symmetric positive definite and the gradient block is minus the
transpose of the divergence block on the free unknowns; the Powell-Hestenes
solve reproduces an independent direct solve of the full saddle-point
system."""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from mdgpu import setups as S


def test_linear_velocity_gives_constant_strain_rates(oracle_mod):
    O = oracle_mod
    g = O.Grid(21, 17, 0.0, 1.0, -0.5, 0.3)
    a, b, c, d = 0.7, -0.4, 0.25, 0.9          # u = a x + b z, v = c x + d z
    xg, zvx = np.asarray(g.xg), np.asarray(g.zvx)
    xvz, zg = np.asarray(g.xvz), np.asarray(g.zg)
    u = (a * xg[None, :] + b * zvx[:, None]).ravel()
    v = (c * xvz[None, :] + d * zg[:, None]).ravel()
    div = a + d
    for comp in (0, 1):
        sr = O.strain_rates(g, u, v, compressible=comp)
        trace = div / 3.0 if comp else 0.0   # the incompressible form keeps the full rates
        np.testing.assert_allclose(sr["div"], div, rtol=0, atol=1e-13)
        np.testing.assert_allclose(sr["exxd"], a - trace, rtol=0, atol=1e-13)
        np.testing.assert_allclose(sr["ezzd"], d - trace, rtol=0, atol=1e-13)
        exz = sr["exz"].reshape(g.Nz, g.Nx)
        np.testing.assert_allclose(exz[1:-1, 1:-1], 0.5 * (b + c), rtol=0, atol=1e-13)


def _picard(O, n=17, seed=3):
    g = O.Grid(n, n, -0.5, 0.5, -0.5, 0.5)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    ph = ((x * x + z * z) < 0.2 ** 2).astype(np.int32)
    eta = np.array([1.0, 1e3])
    en = O.p2mastah(g, x, z, ph, eta, centroid=1)
    es = O.p2mastah(g, x, z, ph, eta, centroid=0)
    num = g.numbering()
    A, B, Cm = O.assemble_picard(g, num, en, es)
    return g, num, en, es, A, B, Cm


def test_picard_operator_structure(oracle_mod):
    g, num, en, es, A, B, Cm = _picard(oracle_mod)
    Ad = A.toarray()
    np.testing.assert_allclose(Ad, Ad.T, rtol=0, atol=1e-10 * np.abs(Ad).max())
    assert np.linalg.eigvalsh(0.5 * (Ad + Ad.T)).min() > 0.0
    np.testing.assert_allclose(B.toarray(), -Cm.toarray().T, rtol=0, atol=1e-12 * np.abs(B.toarray()).max())


def test_killer_solve_matches_a_direct_solve(oracle_mod):
    O = oracle_mod
    from oracle.driver import evaluate_residual, initial_fields
    g, num, en, es, A, B, Cm = _picard(O, n=25)
    nm = num[3]
    u, v, p = initial_fields(g)
    rx, rz = O.eval_rhs(g, 0.0, 0.0, np.zeros(en.size), np.zeros(es.size))
    F, _ = evaluate_residual(g, num, u, v, p, en, es, rx, rz)
    du, dp, st, rc = O.killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], penalty=1e5, max_its_PH=30,
                                    rel_tol_KSP=1e-12, max_its_KSP=500, lin_abs_div=1e-14,
                                    lin_rel_div=1e-14, lin_abs_mom=1e-14, lin_rel_mom=1e-14)
    # direct LU solve of [A B 0; C 0 1; 0 1' 0]: the Lagrange multiplier row
    # fixes the pressure constant the pure-shear box leaves free
    npr = Cm.shape[0]
    one = sp.csr_matrix(np.ones((npr, 1)))
    K = sp.bmat([[A, B, None], [Cm, None, one], [None, one.T, None]]).tocsc()
    x = spla.spsolve(K, np.concatenate([F, [0.0]]))
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(du, x[:nm]) < 1e-7
    xp = x[nm:nm + npr]
    assert rel(dp - dp.mean(), xp - xp.mean()) < 1e-7


def test_killer_solve_incompressibility_abort(oracle_mod):
    """KillerSolver's a-posteriori check (Solvers.c:1333-1336): the solve is
    refused (status -2, the reference exits) when the continuity residual's
    signed maximum stays above 10 lin_abs_div and above lin_rel_div of its
    first Powell-Hestenes value.  One PH iteration with a weak penalty leaves
    a ratio of exactly 1 and a large divergence, so it must abort; the same
    system with the SolVi penalty and enough PH iterations must not."""
    O = oracle_mod
    from oracle.driver import evaluate_residual, initial_fields
    g, num, en, es, A, B, Cm = _picard(O, n=21)
    nm = num[3]
    u, v, p = initial_fields(g)
    rx, rz = O.eval_rhs(g, 0.0, 0.0, np.zeros(en.size), np.zeros(es.size))
    F, _ = evaluate_residual(g, num, u, v, p, en, es, rx, rz)
    _, _, st, rc = O.killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], penalty=1.0, max_its_PH=1)
    assert rc == -2 and st.status == -2
    assert st.maxdiv == st.maxdiv0 and abs(st.maxdiv) > 10 * 1e-9
    _, _, st, rc = O.killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], penalty=1e5, max_its_PH=10)
    assert rc == 0, (st.maxdiv, st.maxdiv0)
    assert abs(st.maxdiv) < 1e-9 or st.maxdiv / st.maxdiv0 < 1e-5
    # the criterion reads the signed maximum, not the largest magnitude
    # (MinMaxArrayVal, MiscFunctions.c:342-375): a negative-only residual
    # has a negative maximum and passes the absolute test whatever its size
    _, _, st, rc = O.killer_solve(g, num, A, B, Cm, F[:nm], -np.abs(F[nm:]) - 1.0, penalty=1.0,
                                  max_its_PH=1)
    assert st.maxdiv <= st.maxdiv0 or st.maxdiv < 0
