"""The oracle's Newton Jacobian (orc_assemble_jacobian, a restatement of
Xjacobian/Zjacobian_InnerNodesDecoupled3 and BuildJacobianOperatorDecoupled,
StokesAssemblyDecoupled.c:219-1085,1786-2304) pinned two ways:

* with the Picard coefficients (D11 = D22 = 2 eta, D33 = eta, every other
  D zero) and comp = 0 it is the Picard assembly itself, boundary rows
  included, for free-slip and no-slip walls (the Picard assembly is pinned
  by SolVi, SolCx and the operator structure tests);
* with arbitrary coefficients, its inner rows are the linearisation of the
  momentum residual with the tangent law
      sxx = D11 exx_d + D12 ezz_d + D13 gxz + D14 p,  szz = D21 ... + D24 p,
      sxz = D31 exx_d + D32 ezz_d + D33 gxz + D34 p,
  exx_d = exx - co (exx + ezz)/3, gxz = 2 exz, centre gxz and vertex
  exx_d, ezz_d, p as means of the four surrounding values — computed here
  directly from the grid, independently of the assembly (co = 0 and 1)."""
import numpy as np
import pytest

from mdgpu import setups as S
from conftest import solvi_bcs


def _case(O, n=19, bc="pure_shear", seed=3):
    g = O.Grid(n, n, -0.5, 0.5, -0.5, 0.5)
    if bc == "pure_shear":
        g.pure_shear(1.0)
    else:
        solvi_bcs(g, S)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    ph = ((x * x + z * z) < 0.2 ** 2).astype(np.int32)
    eta = np.array([1.0, 1e3])
    en = O.p2mastah(g, x, z, ph, eta, centroid=1)
    es = O.p2mastah(g, x, z, ph, eta, centroid=0)
    return g, g.numbering(), en, es


@pytest.mark.parametrize("bc", ["pure_shear", "solvi"])
def test_jacobian_with_picard_coefficients_is_the_picard_operator(oracle_mod, bc):
    O = oracle_mod
    g, num, en, es = _case(O, bc=bc)
    A, B, _ = O.assemble_picard(g, num, en, es)
    Dn = np.zeros((8, en.size)); Ds = np.zeros((4, es.size))
    Dn[0] = Dn[5] = 2 * en
    Ds[2] = es
    AJ, BJ = O.assemble_jacobian(g, num, Dn, Ds, co=0.0)
    assert abs(AJ - A).max() <= 1e-14 * abs(A).max()
    assert abs(BJ - B).max() <= 1e-14 * abs(B).max()


def _linearised_rows(g, num, Dn, Ds, co, u, v, p):
    """Momentum rows of the linearised residual on grid arrays (u, v
    include zero walls; the test only reads rows away from the walls)."""
    Nx, Nz = g.Nx, g.Nz
    dx, dz, dV = g.dx, g.dz, g.dx * g.dz
    U = u.reshape(Nz + 1, Nx)          # [l, k]
    V = v.reshape(Nz, Nx + 1)
    P = p.reshape(Nz - 1, Nx - 1)
    exx = (U[1:-1, 1:] - U[1:-1, :-1]) / dx          # cells [l, k]
    ezz = (V[1:, 1:-1] - V[:-1, 1:-1]) / dz
    div = exx + ezz
    exd, ezd = exx - co * div / 3.0, ezz - co * div / 3.0
    gxz = (U[1:, :] - U[:-1, :]) / dz + (V[:, 1:] - V[:, :-1]) / dx   # vertices [l, k]
    gc = 0.25 * (gxz[:-1, :-1] + gxz[:-1, 1:] + gxz[1:, :-1] + gxz[1:, 1:])
    D = [d.reshape(Nz - 1, Nx - 1) for d in Dn]
    E = [d.reshape(Nz, Nx) for d in Ds]
    sxx = D[0] * exd + D[1] * ezd + D[2] * gc + D[3] * P
    szz = D[4] * exd + D[5] * ezd + D[6] * gc + D[7] * P
    avg = lambda a: 0.25 * (a[:-1, :-1] + a[:-1, 1:] + a[1:, :-1] + a[1:, 1:])   # inner vertices
    sxz = np.zeros((Nz, Nx))
    sxz[1:-1, 1:-1] = (E[0][1:-1, 1:-1] * avg(exd) + E[1][1:-1, 1:-1] * avg(ezd) + E[2][1:-1, 1:-1] * gxz[1:-1, 1:-1]
                       + E[3][1:-1, 1:-1] * avg(P))
    Fx = np.zeros((Nz + 1, Nx))
    Fz = np.zeros((Nz, Nx + 1))
    Fx[1:-1, 1:-1] = -dV * ((sxx[:, 1:] - sxx[:, :-1]) / dx - (P[:, 1:] - P[:, :-1]) / dx
                            + (sxz[1:, 1:-1] - sxz[:-1, 1:-1]) / dz)
    Fz[1:-1, 1:-1] = -dV * ((szz[1:, :] - szz[:-1, :]) / dz - (P[1:, :] - P[:-1, :]) / dz
                            + (sxz[1:-1, 1:] - sxz[1:-1, :-1]) / dx)
    return Fx.ravel(), Fz.ravel()


@pytest.mark.parametrize("co", [0.0, 1.0])
def test_jacobian_inner_rows_linearise_the_tangent_residual(oracle_mod, co):
    O = oracle_mod
    g, num, en, es = _case(O, n=21)
    eu, ev, ep, nm, nc, neq = num
    rng = np.random.default_rng(7)
    Dn = rng.standard_normal((8, en.size))
    Ds = rng.standard_normal((4, es.size))
    AJ, BJ = O.assemble_jacobian(g, num, Dn, Ds, co=co)
    x = rng.standard_normal(nm)
    q = rng.standard_normal(nc)
    y = AJ @ x + BJ @ q
    u = np.zeros(eu.size); v = np.zeros(ev.size); p = np.zeros(ep.size)
    u[eu >= 0] = x[eu[eu >= 0]]
    v[ev >= 0] = x[ev[ev >= 0]]
    p[ep >= 0] = q[ep[ep >= 0] - nm]
    Fx, Fz = _linearised_rows(g, num, Dn, Ds, co, u, v, p)
    Nx, Nz = g.Nx, g.Nz
    checked = 0
    for l in range(3, Nz - 2):
        for k in range(3, Nx - 3):
            e = eu[k + l * Nx]
            assert abs(y[e] - Fx[k + l * Nx]) <= 1e-12 * (abs(Fx).max() + 1.0), ("x", k, l)
            checked += 1
    for l in range(3, Nz - 3):
        for k in range(3, Nx - 2):
            e = ev[k + l * (Nx + 1)]
            assert abs(y[e] - Fz[k + l * (Nx + 1)]) <= 1e-12 * (abs(Fz).max() + 1.0), ("z", k, l)
            checked += 1
    assert checked > 400


def test_jacobian_row_structure(oracle_mod):
    """Inner rows hold the 21-point velocity stencil and 6 pressures of the
    reference (nnzA = 21, nnzB = 6 per row, :1786-2304)."""
    O = oracle_mod
    g, num, en, es = _case(O, n=21)
    rng = np.random.default_rng(2)
    AJ, BJ = O.assemble_jacobian(g, num, rng.standard_normal((8, en.size)), rng.standard_normal((4, es.size)))
    eu, ev = num[0], num[1]
    e = eu[10 + 10 * g.Nx]
    assert AJ.indptr[e + 1] - AJ.indptr[e] == 21 and BJ.indptr[e + 1] - BJ.indptr[e] == 6
    e = ev[10 + 10 * (g.Nx + 1)]
    assert AJ.indptr[e + 1] - AJ.indptr[e] == 21 and BJ.indptr[e + 1] - BJ.indptr[e] == 6
