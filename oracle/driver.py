"""Oracle driver: one defect-correction Stokes step for isotropic viscous
rheology, composed from the restated reference functions.

Order of operations follows the non-linear loop of Main_DOODZ.c:878-1062:
  ApplyBC -> StrainRateComponents -> stresses (2*eta*E, RheologyDensity.c:1311-1313,
  1403-1405) -> EvaluateRHS -> residual F (StokesRoutines.c:867-955) ->
  KillerSolver on J*dx = F (StokesRoutines.c:836-861) -> u <- u - dx
  (ExtractSolutions2 with alpha = -1, StokesRoutines.c:536-576).
TEST INFRASTRUCTURE ONLY.
"""
import numpy as np

from . import (apply_bc, assemble_picard, eval_rhs, killer_solve, residual, strain_rates)


def stresses_viscous(g, u, v, eta_n, eta_s):
    sr = strain_rates(g, u, v)
    return sr, 2.0 * eta_n * sr["exxd"], 2.0 * eta_n * sr["ezzd"], 2.0 * eta_s * sr["exz"]


def evaluate_residual(g, num, u, v, p, eta_n, eta_s, roger_x, roger_z, d=None):
    apply_bc(g, u, v)
    sr, sxx, szz, sxz = stresses_viscous(g, u, v, eta_n, eta_s)
    return residual(g, num, sxx, szz, sxz, p, sr["div"], roger_x, roger_z, d)


def initial_fields(g):
    """Velocity/pressure arrays with Dirichlet values in place (tag 0)."""
    u = np.zeros(g.Nx * (g.Nz + 1)); v = np.zeros((g.Nx + 1) * g.Nz)
    p = np.zeros((g.Nx - 1) * (g.Nz - 1))
    m = g.BCu_t == 0
    u[m] = g.BCu_v[m]
    m = g.BCv_t == 0
    v[m] = g.BCv_v[m]
    apply_bc(g, u, v)
    return u, v, p


def scatter_solution(num, dxv, u, v, p, alpha=-1.0):
    """ExtractSolutions2 (StokesRoutines.c:536-576) for the tags in use."""
    eu, ev, ep, nm, nc, neq = num
    mu = eu >= 0
    u[mu] += alpha * dxv[eu[mu]]
    mv = ev >= 0
    v[mv] += alpha * dxv[ev[mv]]
    mp = ep >= 0
    p[mp] += alpha * dxv[ep[mp]]


def linear_defect_step(g, num, u, v, p, eta_n, eta_s, rho_n, rho_s, gx=0.0, gz=0.0,
                       **killer_kw):
    """One Newton/Picard defect-correction step of a linear viscous problem.
    Returns (norms before, norms after, killer stats)."""
    rx, rz = eval_rhs(g, gx, gz, rho_n, rho_s)
    F, n0 = evaluate_residual(g, num, u, v, p, eta_n, eta_s, rx, rz)
    A, B, Cm = assemble_picard(g, num, eta_n, eta_s)
    nm = num[3]
    du, dp, stats, st = killer_solve(g, num, A, B, Cm, F[:nm], F[nm:], **killer_kw)
    if st != 0:
        raise RuntimeError("KillerSolver status %d" % st)
    scatter_solution(num, np.concatenate([du, dp]), u, v, p)
    apply_bc(g, u, v)
    F1, n1 = evaluate_residual(g, num, u, v, p, eta_n, eta_s, rx, rz)
    return n0, n1, stats, (A, B, Cm, F, F1)
