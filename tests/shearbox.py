"""The visco-elasto-plastic shear-band box used by the GPU-scale tests:
pure shear of a power-law matrix (n = 3) whose viscosity prefactor is
lognormal (sigma = 0.5) per marker and reaches the nodes by P2Mastah's
geometric average, Maxwell elasticity with preloaded old stresses,
Drucker-Prager yield (friction 30 degrees) and 64 seeded weak circles of
radius 2 cells with half the cohesion. Nondimensional: the matrix
viscosity at unit strain rate is 1, G dt = 1 by default; PHYSICAL holds
the scaled values of the physical set-up."""
import numpy as np

import mdgpu
from mdgpu import setups as S
from mdgpu.nonlinear import PicardLoop


# The shear-band benchmark's physical parameters, scaled as the reference
# scales its inputs (ScaleMe, InputOutput.c:1835-1849) with a viscosity
# scale of 1e22 Pa s and a strain-rate scale of 1e-15 1/s, i.e. a stress
# scale of 1e7 Pa and a time scale of 1e15 s: shear modulus 1e10 Pa,
# cohesion 1e7 Pa (seeds 5e6 Pa), friction 30 degrees, dt = 1e11 s,
# background pure shear 1e-15 1/s, matrix viscosity 1e22 Pa s at that rate.
PHYSICAL = dict(G=1e10 / 1e7, C=1e7 / 1e7, dt=1e11 / 1e15)


def build(O, n, C, device=False, G=1.0, dt=1.0, pre=None):
    """Returns (ctx, loop, st, info) for an n x n box with cohesion C; device
    selects the loop that keeps its fields in device memory.  G and dt are
    the (scaled) shear modulus and time step; the old deviatoric stresses
    are 0.2 with a normal jitter of 0.05, or pre * C with a 5% jitter when
    pre is given (a box loaded close to yield)."""
    from oracle.driver import initial_fields
    box = (0.0, 1.0, -1.0, 0.0)
    g = O.Grid(n + 1, n + 1, *box)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=7)
    ph = S.seed_phase(x, z, S.weak_seeds(9, 64, 2 * g.dx, *box)).astype(np.int32)
    m = np.arange(x.size, dtype=np.uint64)
    u1, u2 = S.uniform(77, 2 * m), S.uniform(77, 2 * m + np.uint64(1))
    pref_m = np.exp(0.5 * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2))   # Box-Muller
    ctx = mdgpu.Context(n + 1, n + 1, *box)
    ctx.set_bc(g.BCu_t, g.BCu_v, g.BCv_t, g.BCv_v, g.BCp_t, g.BCg_t)
    ctx.markers_set(x, z, ph)
    perc_n = np.array(ctx.p2mastah(np.zeros(2), prop=1, centroid=1, nb_phases=2))
    perc_s = np.array(ctx.p2mastah(np.zeros(2), prop=1, centroid=0, nb_phases=2))
    pref_n = ctx.p2mastah(pref_m, flag=1, avg=2, centroid=1)
    pref_s = ctx.p2mastah(pref_m, flag=1, avg=2, centroid=0)
    nc, nv = n * n, (n + 1) ** 2
    u, v, p = initial_fields(g)
    noise = np.random.default_rng(3).standard_normal(nc)
    s0 = 0.2 + 0.05 * noise if pre is None else pre * C * (1.0 + 0.05 * noise)
    st = dict(u=u, v=v, p=p, T=np.ones(nc), mu_n=np.full(nc, G), mu_s=np.full(nv, G), perc_n=perc_n,
              perc_s=perc_s, sxxd0=s0, szzd0=-s0, sxz0_n=np.zeros(nc), sxxd0_s=ctx.c2v(s0),
              szzd0_s=ctx.c2v(-s0), sxz0=np.zeros(nv), rho_n=np.ones(nc), rho_s=np.ones(nv),
              pref_n=pref_n, pref_s=pref_s)
    prm = mdgpu.RheoParams(average=0, step=1, elastic=1, dt=dt, R=1.0, zeroC=0.0, min_eta=1e-4,
                           max_eta=1e4, eta_tol=1e-13, Tiimin=1e-5, scale_E=1.0, scale_eta=1.0)
    eu, ev, ep = g.numbering()[:3]
    loop = PicardLoop(ctx, eu, ev, ep, [mdgpu.PhaseProp(**d) for d in S.vep_phases(C=C)], prm, device=device)
    info = dict(markers=x.size, weak=int(ph.sum()), pref_m=pref_m, nc=nc)
    return ctx, loop, st, info
