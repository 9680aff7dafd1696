"""Host driver of the learned preconditioner: problems, pair capture, training.

The network and its training run on the GPU (mdg_pcnet_* entry points); this
module only sequences the calls:

1. build seeded random-ellipse viscosity fields on the grid (markers, GPU P2G
   with geometric averaging, GPU body force and residual at u = p = 0);
2. solve them with the plain method (FGCR + multigrid CG, precond = 0) while
   the library captures the (right-hand side, solution) pairs of the inner
   velocity solves;
3. train the network on those pairs (Adam, on the GPU);
4. solve with precond = 1: every inner multigrid CG starts from the
   network's guess, and the guess is dropped whenever it does not lower the
   residual, so the plain method remains the fallback and decides when the
   solve has converged.

The outer variant (collect_outer_pairs, precond = 2) captures instead the
(residual R_k, error x* - x_k) pairs of the outer FGCR iterations and, once
trained, adds the network's velocity and pressure directions for the current
residual to every outer iteration; GCR takes each with its own optimal
coefficient, so a poor direction costs an operator application and no
accuracy.
"""
import numpy as np

from . import MDG_SRC_SORTED
from . import workloads as W

# (Cin, Cout) of the seven 3x3 convolutions, in the library's parameter order
LAYERS = [(4, 32), (32, 32), (32, 64), (64, 64), (64, 32), (32, 32), (32, 3)]


def init_params(seed):
    """He-normal weights, zero biases, flattened as W[tap][ci][co] then b."""
    rng = np.random.default_rng(seed)
    out = []
    for ci, co in LAYERS:
        out.append((rng.standard_normal(9 * ci * co) * np.sqrt(2.0 / (9 * ci))).astype(np.float32))
        out.append(np.zeros(co, np.float32))
    return np.concatenate(out)


def free_slip_bc(Nx):
    """Tags of a closed free-slip box (walls 0, ghosts 13), values 0."""
    Nz = Nx
    BCu_t = np.full(Nx * (Nz + 1), -1, np.int8)
    BCv_t = np.full((Nx + 1) * Nz, -1, np.int8)
    for l in range(Nz + 1):
        for k in range(Nx):
            c = k + l * Nx
            if l == 0 or l == Nz:
                BCu_t[c] = 13
            elif k == 0 or k == Nx - 1:
                BCu_t[c] = 0
    for l in range(Nz):
        for k in range(Nx + 1):
            c = k + l * (Nx + 1)
            if k == 0 or k == Nx:
                BCv_t[c] = 13
            elif l == 0 or l == Nz - 1:
                BCv_t[c] = 0
    return (BCu_t, np.zeros(BCu_t.size), BCv_t, np.zeros(BCv_t.size),
            np.full((Nx - 1) * (Nz - 1), -1, np.int8), np.full(Nx * Nz, -1, np.int8))


class EllipseProblem:
    """One seeded random-ellipse Stokes problem on a context's grid."""

    def __init__(self, ctx, ncells, seed, n_ell, nper=3, gz=-9.81):
        wl = W.ellipse_workload(ncells, nper, seed, n_ell=n_ell)
        n = wl["x"].size
        Nx = wl["Nx"]
        nc, nv = (Nx - 1) ** 2, Nx * Nx
        ctx.markers_set(wl["x"], wl["z"], wl["phase"])
        es_m = ctx.device(n)
        rs_m = ctx.device(n)
        ctx.markers_gather(wl["eta_table"][wl["phase"]], es_m)
        ctx.markers_gather(wl["rho_table"][wl["phase"]], rs_m)
        srt = 1 | MDG_SRC_SORTED
        self.eta_n, self.rho_n = np.zeros(nc), np.zeros(nc)
        self.eta_s, self.rho_s = np.zeros(nv), np.zeros(nv)
        ctx.p2mastah_fused(1, 1, [dict(src=es_m, dst=self.eta_n, flag=srt, avg=2),
                                  dict(src=rs_m, dst=self.rho_n, flag=srt, avg=0)])
        ctx.p2mastah_fused(0, 1, [dict(src=es_m, dst=self.eta_s, flag=srt, avg=2),
                                  dict(src=rs_m, dst=self.rho_s, flag=srt, avg=0)])
        es_m.free()
        rs_m.free()
        rx, rz = ctx.eval_rhs(0.0, gz, self.rho_n, self.rho_s)
        z = np.zeros(nc)
        self.F, _ = ctx.stokes_residual(z, z, np.zeros(nv), z, z, rx, rz)
        self.nm = ctx.eqn_count()[1]

    def solve(self, ctx, **opts):
        D11 = 2.0 * self.eta_n
        return ctx.killer_solver(D11, D11, self.eta_s, self.F[:self.nm], self.F[self.nm:], **opts)


def collect_pairs(ctx, problems, max_pairs, **opts):
    """Plain solves of the problems with pair capture on; returns the stats."""
    ctx.pcnet_capture(max_pairs)
    stats = [p.solve(ctx, precond=0, **opts)[1] for p in problems]
    ctx.pcnet_capture(0)            # stop capturing, keep the pairs
    return stats, ctx.pcnet_captured()[0]


def train(ctx, steps, batch=4, lr=1e-3, seed=0, init_seed=1):
    ctx.pcnet_set_params(init_params(init_seed))
    return ctx.pcnet_train(steps, batch=batch, lr=lr, seed=seed)


def train_residual(ctx, steps, batch=4, lr=1e-3, seed=0, init_seed=1, warm_steps=0):
    """Residual-loss training on outer pairs (precond = 2): optionally
    warm-started by warm_steps of the error (L2) loss on the same pairs."""
    ctx.pcnet_set_params(init_params(init_seed))
    if warm_steps:
        ctx.pcnet_train(warm_steps, batch=batch, lr=lr, seed=seed + 1)
    return ctx.pcnet_train_residual(steps, batch=batch, lr=lr, seed=seed)


def collect_outer_pairs(ctx, problems, max_pairs, **opts):
    """Plain solves with outer (residual, error) pair capture on."""
    ctx.pcnet_capture_outer(max_pairs)
    stats = [p.solve(ctx, precond=0, **opts)[1] for p in problems]
    ctx.pcnet_capture(0)
    return stats, ctx.pcnet_captured()[0]
