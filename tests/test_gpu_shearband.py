"""The visco-elasto-plastic shear-band box at full size, on the GPU only.

1023 x 1023 cells, 16 markers per cell (16.7M markers), a power-law matrix
(n = 3) whose viscosity prefactor is lognormal (sigma = 0.5) per marker and
reaches the nodes by P2Mastah's geometric average, Maxwell elasticity with
preloaded old stresses, Drucker-Prager yield (friction 30 degrees) and 64
seeded weak circles of radius 2 cells with half the cohesion, pure shear,
10 non-linear iterations (Picard with the line search of
StokesRoutines.c:582-807) toward a 1e-8 tolerance. Quantities are
nondimensional (the matrix viscosity at unit strain rate is 1, G dt = 1,
C = 1.5: the matrix sits below yield away from the seeds, so the yielding
region is the bands that grow from them).

The oracle cannot run at this size, so the test checks what the size does
not change: phase fractions sum to one on every node; the nodal prefactor is
a geometric mean of the marker values (inside their range, log-mean near
0); every linear solve reaches its tolerance; the continuity residual ends
at round-off; the momentum residuals fall by more than 10x; plastic strain
rates appear on a localised part of the box (more than the seeds, well
under half of it). The loop runs with its fields in device memory (uploaded
once); the wall time of each iteration is printed. The same loop is checked
against the oracle at 31^2 in test_gpu_nonlinear.py, and the device-resident
loop against the host-staged one there too.
"""
import time

import numpy as np
import pytest

from shearbox import PHYSICAL, build

pytestmark = pytest.mark.gpu


def test_shear_band_box_1023(oracle_mod):
    t0 = time.time()
    ctx, loop, st, info = build(oracle_mod, 1023, C=1.5, device=True)
    nc, pref_m = info["nc"], info["pref_m"]
    print("markers", info["markers"], "in weak seeds", info["weak"], "set-up %.1fs" % (time.time() - t0),
          flush=True)
    assert np.abs(st["perc_n"].sum(axis=0) - 1.0).max() < 1e-12
    assert np.abs(st["perc_s"].sum(axis=0) - 1.0).max() < 1e-12
    for f in (st["pref_n"], st["pref_s"]):
        assert f.min() >= pref_m.min() and f.max() <= pref_m.max()
        assert abs(np.log(f).mean()) < 0.01
    t0 = time.time()
    hist, nsolve, stats = loop.run(st, 10, (1e-12, 1e-8, 1e-12, 1e-8), diag_scaling=1,
                                   line_search=True, rel_tol=1e-8)
    wall = time.time() - t0
    h = np.array(hist)
    rel = h / h[0]
    print("linear solves", nsolve, "in %.1fs" % wall, "(upload %.0f ms)" % st["upload_ms"], "step lengths",
          st["alphas"], flush=True)
    print("ms per iteration", [round(t) for t in st["iter_ms"]], "solve ms", [round(o.time_ms) for o in stats],
          flush=True)
    print("outer its", [o.its for o in stats], "inner its", [o.inner_its for o in stats], flush=True)
    print("relative residual history (x, z, continuity)\n", rel[:, :3], flush=True)
    assert nsolve == len(stats) >= 10
    assert all(o.rel_res <= 1e-8 for o in stats)
    assert rel[-1, 0] < 0.1 and rel[-1, 1] < 0.1
    assert rel[-1, 2] < 1e-8
    assert st["diverged_solves"] == []
    loop.device = False     # one host-staged evaluation at the final state
    _, cen, ver = loop.update_nonlinearity(st)
    npl = int((cen["eII_pl"] > 0).sum())
    print("centres yielding", npl, "of", nc, flush=True)
    seeds_cells = 64 * np.pi * 2.0 ** 2
    assert seeds_cells < npl < 0.5 * nc


def test_shear_band_box_1023_physical_newton(oracle_mod):
    """The box with the physical parameters (shear modulus 1e10 Pa, cohesion
    1e7 Pa with 5e6 Pa in the seeds, friction 30 degrees, dt = 1e11 s,
    pure shear at 1e-15 1/s, matrix viscosity 1e22 Pa s at that rate; scaled
    by 1e22 Pa s and 1e-15 1/s as the reference scales its inputs), loaded
    close to yield (old deviatoric stresses 0.8 C): Picard iterations with
    the line search switch to Newton once the residual fell tenfold
    (Main_DOODZ.c:884-896), and the loop reaches the 1e-8 non-linear
    tolerance within its 10 iterations, the Newton steps converging
    quadratically; the solver runs the vertex-patch smoother (penalty 100,
    degree 12 on every level) and every solve meets 1e-8."""
    ctx, loop, st, info = build(oracle_mod, 1023, device=True, pre=0.8, **PHYSICAL)
    t0 = time.time()
    hist, nsolve, stats = loop.run(st, 10, (1e-12, 1e-8, 1e-12, 1e-8), diag_scaling=1, line_search=True,
                                   newton=True, smoother=3, penalty=100.0, cheb_degree=12, cheb_degree_coarse=12)
    h = np.array(hist)
    rel = h / h[0]
    print("solves", nsolve, "in %.1fs" % (time.time() - t0), "Newton steps", st["newton_steps"], "alphas",
          st["alphas"], flush=True)
    print("ms per iteration", [round(t) for t in st["iter_ms"]], "outer its", [o.its for o in stats], flush=True)
    print("relative residual history (x, z, continuity)\n", rel[:, :3], flush=True)
    assert st["newton_steps"] and nsolve <= 10
    assert all(o.rel_res <= 1e-8 for o in stats)
    assert rel[-1, 0] < 1e-8 and rel[-1, 1] < 1e-8 and rel[-1, 2] < 1e-8
    # quadratic convergence of the last Newton steps
    assert rel[-1, 0] < 1e-2 * rel[-2, 0]
