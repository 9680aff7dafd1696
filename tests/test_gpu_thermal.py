"""GPU energy step (mdg_energy_step) against the oracle's restatement of
EnergyDirectSolve + SolveThermalPCG (ThermalRoutines.c:74-470,
ThermalSolver.c:338-480).  Each cell's matrix row and right-hand side
follow the reference's order of operations, so the PCG iterates differ only
through the order of the dot-product sums: iteration counts agree to +-1
(a count lands on the other side of the tolerance only when the ratio sits
within round-off of it) and temperatures to 1e-10 of their range.  Cases:
heterogeneous conductivities with all source terms and both boundary
types, the reference's GaussianDiffusionL2 known answer, the iteration
limit (-1), a zero initial residual (0), device-resident inputs, and the
2047^2 grid."""
import time

import numpy as np
import pytest

import mdgpu

pytestmark = pytest.mark.gpu


def _case(O, ncx, ncz, seed, L=(3.0, 2.0)):
    g = O.Grid(ncx + 1, ncz + 1, 0.0, L[0], -L[1], 0.0)
    rng = np.random.default_rng(seed)
    nc = ncx * ncz
    d = {
        "rho_n": rng.uniform(2.5, 3.5, nc), "Cp": rng.uniform(900.0, 1100.0, nc),
        "T": rng.uniform(300.0, 1600.0, nc),
        "kx": rng.uniform(1.0, 5.0, g.Nx * (g.Nz + 1)), "kz": rng.uniform(1.0, 5.0, (g.Nx + 1) * g.Nz),
        "Qr": rng.uniform(0.0, 1e3, nc), "Wdiss": rng.uniform(0.0, 1e3, nc), "Ha": rng.uniform(-1e2, 1e2, nc),
    }
    BT = np.zeros((ncz + 2, ncx + 2), np.int8)
    BV = np.zeros((ncz + 2, ncx + 2))
    BT[-1, :] = 1; BV[-1, :] = 273.0          # top: Dirichlet
    BT[0, :] = 0; BV[0, :] = 0.05             # bottom: insulated + flux term
    BT[:, 0] = 1; BV[:, 0] = 900.0            # west wall: Dirichlet
    d["BCT_t"], d["BCT_v"] = BT.ravel(), BV.ravel()
    return g, d


def _both(O, g, d, dt, max_its, tol, device=False, **drop):
    args = {k: v for k, v in d.items() if k not in drop}
    T_o, its_o = O.energy_step(g, dt, args["rho_n"], args["Cp"], args["T"], args["kx"], args["kz"],
                               args["BCT_t"], args["BCT_v"], Qr=args.get("Qr"), Wdiss=args.get("Wdiss"),
                               Ha=args.get("Ha"), max_its=max_its, rel_tol=tol)
    ctx = mdgpu.Context(g.Nx, g.Nz, g.xg[0], g.xg[-1], g.zg[0], g.zg[-1])
    T = args["T"].copy()
    if device:
        dev = {k: mdgpu.DeviceArray(ctx, v.size, v.dtype, host=v) for k, v in args.items()}
        its = ctx.energy_step(dt, dev["rho_n"], dev["Cp"], dev["T"], dev["kx"], dev["kz"], dev["BCT_t"],
                              dev["BCT_v"], Qr=dev.get("Qr"), Wdiss=dev.get("Wdiss"), Ha=dev.get("Ha"),
                              max_its=max_its, rel_tol=tol)
        T = dev["T"].numpy()
    else:
        its = ctx.energy_step(dt, args["rho_n"], args["Cp"], T, args["kx"], args["kz"], args["BCT_t"],
                              args["BCT_v"], Qr=args.get("Qr"), Wdiss=args.get("Wdiss"), Ha=args.get("Ha"),
                              max_its=max_its, rel_tol=tol)
    return T, its, T_o, its_o, ctx


def _close(T, T_o, tol=1e-10):
    scale = max(np.ptp(T_o), np.abs(T_o).max() * 1e-3)
    assert np.abs(T - T_o).max() <= tol * scale, np.abs(T - T_o).max() / scale


@pytest.mark.parametrize("shape", [(40, 30), (97, 61)])
@pytest.mark.parametrize("device", [False, True])
def test_energy_step_matches_oracle(oracle_mod, shape, device):
    g, d = _case(oracle_mod, *shape, seed=shape[0])
    T, its, T_o, its_o, _ = _both(oracle_mod, g, d, 0.05, 2000, 1e-12, device=device)
    assert its_o > 5
    assert abs(its - its_o) <= 1
    _close(T, T_o)


def test_optional_sources_absent(oracle_mod):
    g, d = _case(oracle_mod, 33, 47, seed=3)
    T, its, T_o, its_o, _ = _both(oracle_mod, g, d, 0.2, 2000, 1e-11, Wdiss=1, Ha=1)
    assert abs(its - its_o) <= 1
    _close(T, T_o)


def test_iteration_limit_and_zero_residual(oracle_mod):
    g, d = _case(oracle_mod, 50, 40, seed=9)
    T, its, T_o, its_o, _ = _both(oracle_mod, g, d, 0.05, 3, 1e-14)
    assert its == its_o == -1
    _close(T, T_o, 1e-12)
    # zero sources, zero boundary values, T = 0: b = 0 = A x, nothing to do
    z = dict(d, T=np.zeros_like(d["T"]), BCT_v=np.zeros_like(d["BCT_v"]))
    T, its, T_o, its_o, _ = _both(oracle_mod, g, z, 0.05, 50, 1e-8, Qr=1, Wdiss=1, Ha=1)
    assert its == its_o == 0
    assert np.all(T == 0.0)


def test_gaussian_diffusion_known_answer(oracle_mod):
    """TESTS/ThermalTests.cpp:191-229 (GaussianDiffusionL2): five steps of
    1e12 s on the 51 x 51 grid, relative L2 error against the analytic
    Gaussian below the reference's 2e-2 (the grid step alone: below 1e-3)."""
    O = oracle_mod
    n = 51
    g = O.Grid(n, n, -5e4, 5e4, -5e4, 5e4)
    nc = (n - 1) ** 2
    XC, ZC = np.meshgrid(g.xc, g.zc)
    Tbg, dT, R = 773.15, 200.0, 1e4
    T = (Tbg + dT * np.exp(-(XC ** 2 + ZC ** 2) / R ** 2)).ravel()
    T_o = T.copy()
    BT = np.zeros((n + 1) ** 2, np.int8)
    BV = np.zeros(BT.size)
    kx, kz = np.full(g.Nx * (g.Nz + 1), 3.0), np.full((g.Nx + 1) * g.Nz, 3.0)
    rho, cp = np.full(nc, 3300.0), np.full(nc, 1050.0)
    ctx = mdgpu.Context(n, n, -5e4, 5e4, -5e4, 5e4)
    for _ in range(5):
        its = ctx.energy_step(1e12, rho, cp, T, kx, kz, BT, BV, rel_tol=1e-12, max_its=1000)
        T_o, its_o = O.energy_step(g, 1e12, rho, cp, T_o, kx, kz, BT, BV, rel_tol=1e-12, max_its=1000)
        assert its > 0 and abs(its - its_o) <= 1
    _close(T, T_o)
    R2 = R * R + 4.0 * 3.0 / (3300.0 * 1050.0) * 5e12
    Ta = (Tbg + dT * R * R / R2 * np.exp(-(XC ** 2 + ZC ** 2) / R2)).ravel()
    assert np.sqrt(((T - Ta) ** 2).sum() / (Ta ** 2).sum()) < 1e-3


def test_energy_step_2047(oracle_mod):
    """The 2047^2 grid (4.19M cells): parity with the oracle on the whole
    field, and the PCG's time per iteration printed."""
    g, d = _case(oracle_mod, 2046, 2046, seed=1, L=(1.0, 1.0))
    t0 = time.time()
    T, its, T_o, its_o, ctx = _both(oracle_mod, g, d, 0.01, 3000, 1e-10)
    print("2047^2 energy step: %d its (oracle %d), GPU %.1f ms (%.3f ms/it), wall %.1f s"
          % (its, its_o, ctx.last_kernel_ms(), ctx.last_kernel_ms() / max(its, 1), time.time() - t0))
    assert its > 0 and abs(its - its_o) <= 1
    _close(T, T_o, 1e-9)
