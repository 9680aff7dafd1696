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
under half of it). The same loop is checked against the oracle at 31^2 in
test_gpu_nonlinear.py.
"""
import time

import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S
from mdgpu.nonlinear import PicardLoop

pytestmark = pytest.mark.gpu


def test_shear_band_box_1023(oracle_mod):
    from oracle.driver import initial_fields
    O = oracle_mod
    n = 1023
    box = (0.0, 1.0, -1.0, 0.0)
    t0 = time.time()
    g = O.Grid(n + 1, n + 1, *box)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=7)
    ph = S.seed_phase(x, z, S.weak_seeds(9, 64, 2 * g.dx, *box)).astype(np.int32)
    m = np.arange(x.size, dtype=np.uint64)
    u1, u2 = S.uniform(77, 2 * m), S.uniform(77, 2 * m + np.uint64(1))
    pref_m = np.exp(0.5 * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2))   # Box-Muller
    print("markers", x.size, "in weak seeds", int(ph.sum()), "set-up %.1fs" % (time.time() - t0), flush=True)

    ctx = mdgpu.Context(n + 1, n + 1, *box)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    ctx.markers_set(x, z, ph)
    perc_n = np.array(ctx.p2mastah(np.zeros(2), prop=1, centroid=1, nb_phases=2))
    perc_s = np.array(ctx.p2mastah(np.zeros(2), prop=1, centroid=0, nb_phases=2))
    pref_n = ctx.p2mastah(pref_m, flag=1, avg=2, centroid=1)
    pref_s = ctx.p2mastah(pref_m, flag=1, avg=2, centroid=0)
    assert np.abs(perc_n.sum(axis=0) - 1.0).max() < 1e-12
    assert np.abs(perc_s.sum(axis=0) - 1.0).max() < 1e-12
    for f in (pref_n, pref_s):
        assert f.min() >= pref_m.min() and f.max() <= pref_m.max()
        assert abs(np.log(f).mean()) < 0.01

    nc, nv = n * n, (n + 1) ** 2
    u, v, p = initial_fields(g)
    s0 = 0.2 + 0.05 * np.random.default_rng(3).standard_normal(nc)
    st = dict(u=u, v=v, p=p, T=np.ones(nc), mu_n=np.ones(nc), mu_s=np.ones(nv), perc_n=perc_n,
              perc_s=perc_s, sxxd0=s0, szzd0=-s0, sxz0_n=np.zeros(nc), sxxd0_s=ctx.c2v(s0),
              szzd0_s=ctx.c2v(-s0), sxz0=np.zeros(nv), rho_n=np.ones(nc), rho_s=np.ones(nv),
              pref_n=pref_n, pref_s=pref_s)
    prm = mdgpu.RheoParams(average=0, step=1, elastic=1, dt=1.0, R=1.0, zeroC=0.0, min_eta=1e-4,
                           max_eta=1e4, eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in S.vep_phases(C=1.5)], prm)
    t0 = time.time()
    hist, nsolve, stats = loop.run(st, 10, (1e-12, 1e-8, 1e-12, 1e-8), diag_scaling=1,
                                   line_search=True, rel_tol=1e-8)
    wall = time.time() - t0
    h = np.array(hist)
    rel = h / h[0]
    print("linear solves", nsolve, "in %.1fs" % wall, "step lengths", st["alphas"], flush=True)
    print("outer its", [o.its for o in stats], "inner its", [o.inner_its for o in stats], flush=True)
    print("relative residual history (x, z, continuity)\n", rel[:, :3], flush=True)
    assert nsolve == len(stats) >= 10
    assert all(o.rel_res <= 1e-8 for o in stats)
    assert rel[-1, 0] < 0.1 and rel[-1, 1] < 0.1
    assert rel[-1, 2] < 1e-8
    _, cen, ver = loop.update_nonlinearity(st)
    npl = int((cen["eII_pl"] > 0).sum())
    print("centres yielding", npl, "of", nc, flush=True)
    seeds_cells = 64 * np.pi * 2.0 ** 2
    assert seeds_cells < npl < 0.5 * nc
