"""Oracle pins of the implicit energy step (EnergyDirectSolve,
ThermalRoutines.c:74-470, with SolveThermalPCG, ThermalSolver.c:338-480),
from what the mathematics fixes:

* with constant coefficients a separable cosine (insulated walls) or sine
  (Dirichlet T = 0 at the wall faces) mode is an eigenvector of the
  cell-centred 5-point operator with eigenvalue (4/h^2) sin^2(pi m h / 2L)
  per direction, so one step multiplies it by 1 / (1 + dt kappa (lx + lz));
  a missing factor 2 in the Dirichlet ghost term or a transposed conductivity
  breaks it;
* a linear profile between two Dirichlet walls is reproduced exactly by the
  ghost-node Dirichlet condition, so a long step from T = 0 converges to it;
* insulated walls conserve the heat content: sum(rho Cp T) grows by exactly
  dt * sum(Qr);
* a conductivity layered in z carries the same discrete flux through every
  horizontal face at steady state (pins which Vz node each face uses);
* the matrix is symmetric (the reference solves it with CG)."""
import numpy as np
import pytest


def _setup(O, ncx=24, ncz=16, L=(2.0, 1.0)):
    g = O.Grid(ncx + 1, ncz + 1, 0.0, L[0], -L[1], 0.0)
    nc = ncx * ncz
    BCT_t = np.zeros((ncz + 2) * (ncx + 2), np.int8)
    BCT_v = np.zeros((ncz + 2) * (ncx + 2))
    kx = np.ones(g.Nx * (g.Nz + 1))
    kz = np.ones((g.Nx + 1) * g.Nz)
    return g, nc, BCT_t, BCT_v, kx, kz


def _eig(m, n):
    return 4.0 * np.sin(np.pi * m / (2.0 * n)) ** 2


@pytest.mark.parametrize("dirichlet", [False, True])
def test_mode_decay(oracle_mod, dirichlet):
    O = oracle_mod
    ncx, ncz = 24, 16
    g, nc, BT, BV, kx, kz = _setup(O, ncx, ncz)
    kap, rho, cp = 2.5, 3.0, 1.5
    kx[:] = kap
    kz[:] = kap
    mx, mz = 3, 2
    k = (np.arange(ncx) + 0.5) / ncx
    l = (np.arange(ncz) + 0.5) / ncz
    if dirichlet:
        f = np.outer(np.sin(np.pi * mz * l), np.sin(np.pi * mx * k))
        BT2 = BT.reshape(ncz + 2, ncx + 2)
        BT2[0, :] = 1; BT2[-1, :] = 1; BT2[:, 0] = 1; BT2[:, -1] = 1
    else:
        f = np.outer(np.cos(np.pi * mz * l), np.cos(np.pi * mx * k))
    T0 = f.ravel().copy()
    dt = 0.01
    T1, its = O.energy_step(g, dt, np.full(nc, rho), np.full(nc, cp), T0, kx, kz, BT, BV, rel_tol=1e-14,
                            max_its=2000)
    assert its > 0
    lam = _eig(mx, ncx) / g.dx ** 2 + _eig(mz, ncz) / g.dz ** 2
    factor = 1.0 / (1.0 + dt * kap / (rho * cp) * lam)
    np.testing.assert_allclose(T1, factor * T0, rtol=0, atol=1e-12)
    assert factor < 0.9


def test_linear_steady_profile(oracle_mod):
    O = oracle_mod
    ncx, ncz = 8, 20
    g, nc, BT, BV, kx, kz = _setup(O, ncx, ncz, L=(1.0, 1.0))
    BT2, BV2 = BT.reshape(ncz + 2, ncx + 2), BV.reshape(ncz + 2, ncx + 2)
    BT2[0, :] = 1; BV2[0, :] = 1300.0     # bottom (z = -1)
    BT2[-1, :] = 1; BV2[-1, :] = 0.0      # top (z = 0)
    T, its = O.energy_step(g, 1e14, np.ones(nc), np.ones(nc), np.zeros(nc), kx, kz, BT, BV, rel_tol=1e-15,
                           max_its=5000)
    ZC = np.repeat(g.zc, ncx)
    np.testing.assert_allclose(T, -1300.0 * ZC, rtol=0, atol=1e-8)


def test_insulated_heat_content(oracle_mod):
    O = oracle_mod
    ncx, ncz = 20, 14
    g, nc, BT, BV, kx, kz = _setup(O, ncx, ncz)
    rng = np.random.default_rng(5)
    kx[:] = rng.uniform(0.5, 4.0, kx.size)
    kz[:] = rng.uniform(0.5, 4.0, kz.size)
    rho, cp = rng.uniform(2.0, 3.0, nc), rng.uniform(800.0, 1200.0, nc)
    T0, Qr = rng.uniform(0.0, 1000.0, nc), rng.uniform(0.0, 50.0, nc)
    dt = 3.0
    T1, its = O.energy_step(g, dt, rho, cp, T0, kx, kz, BT, BV, Qr=Qr, rel_tol=1e-15, max_its=5000)
    assert its > 0
    lhs = np.sum(rho * cp * T1)
    rhs = np.sum(rho * cp * T0) + dt * Qr.sum()
    assert abs(lhs - rhs) < 1e-12 * rhs
    # the reference's south-wall flux term for an insulated type-0 ghost
    # adds -val/dx to every bottom-row cell's right-hand side
    BV.reshape(ncz + 2, ncx + 2)[0, 1:-1] = 7.0
    T2, _ = O.energy_step(g, dt, rho, cp, T0, kx, kz, BT, BV, Qr=Qr, rel_tol=1e-15, max_its=5000)
    rhs2 = rhs + dt * ncx * (-7.0 / g.dx)
    assert abs(np.sum(rho * cp * T2) - rhs2) < 1e-12 * abs(rhs2)


def test_layered_flux_and_symmetry(oracle_mod):
    O = oracle_mod
    ncx, ncz = 6, 30
    g, nc, BT, BV, kx, kz = _setup(O, ncx, ncz, L=(1.0, 1.0))
    kz2 = kz.reshape(g.Nz, g.Nx + 1)
    kz2[:] = np.linspace(1.0, 9.0, g.Nz)[:, None] ** 1.5
    BT2, BV2 = BT.reshape(ncz + 2, ncx + 2), BV.reshape(ncz + 2, ncx + 2)
    BT2[0, :] = 1; BV2[0, :] = 1.0
    BT2[-1, :] = 1; BV2[-1, :] = 0.0
    T, _ = O.energy_step(g, 1e14, np.ones(nc), np.ones(nc), np.zeros(nc), kx, kz, BT, BV, rel_tol=1e-15,
                         max_its=5000)
    T2 = T.reshape(ncz, ncx)
    # interior faces l | l+1 sit on Vz row l+1; the walls use half a cell
    q_in = -kz2[1:-1, 1:-1] * (T2[1:] - T2[:-1]) / g.dz
    q_bot = -kz2[0, 1:-1] * (T2[0] - 1.0) / (0.5 * g.dz)
    q_top = -kz2[-1, 1:-1] * (0.0 - T2[-1]) / (0.5 * g.dz)
    q = np.vstack([q_bot[None], q_in, q_top[None]])
    np.testing.assert_allclose(q, q[0, 0], rtol=1e-9)
    A, _ = O.energy_matrix(g, 0.5, np.ones(nc), np.ones(nc), np.random.default_rng(1).uniform(1, 2, kx.size), kz,
                           BT, BV)
    assert abs(A - A.T).max() == 0.0


def test_gaussian_diffusion_known_answer(oracle_mod):
    """The reference's GaussianDiffusionL2 case (TESTS/ThermalTests.cpp:191-229,
    TESTS/Thermal/GaussianDiffusionL2.txt): 51 x 51 vertices on +-50 km,
    k = 3, rho = 3300, Cp = 1050, insulated walls, five steps of 1e12 s from
    773.15 K + 200 K exp(-r^2 / R^2), R = 10 km, against the analytic
    Gaussian; the reference accepts a relative L2 error below 2e-2 (its
    marker round trip included).  The grid step alone stays within 1e-3."""
    O = oracle_mod
    n = 51
    g = O.Grid(n, n, -5e4, 5e4, -5e4, 5e4)
    ncx = ncz = n - 1
    nc = ncx * ncz
    XC, ZC = np.meshgrid(g.xc, g.zc)
    Tbg, dT, R = 773.15, 200.0, 1e4
    T = (Tbg + dT * np.exp(-(XC ** 2 + ZC ** 2) / R ** 2)).ravel()
    BT = np.zeros((ncz + 2) * (ncx + 2), np.int8)
    BV = np.zeros(BT.size)
    kx, kz = np.full(g.Nx * (g.Nz + 1), 3.0), np.full((g.Nx + 1) * g.Nz, 3.0)
    rho, cp = np.full(nc, 3300.0), np.full(nc, 1050.0)
    for _ in range(5):
        T, its = O.energy_step(g, 1e12, rho, cp, T, kx, kz, BT, BV, rel_tol=1e-12, max_its=1000)
        assert its > 0
    R2 = R * R + 4.0 * 3.0 / (3300.0 * 1050.0) * 5e12
    Ta = (Tbg + dT * R * R / R2 * np.exp(-(XC ** 2 + ZC ** 2) / R2)).ravel()
    L2 = np.sqrt(((T - Ta) ** 2).sum() / (Ta ** 2).sum())
    assert L2 < 2e-2
    assert L2 < 1e-3
