"""The learned part end to end on the GPU: the plain solver generates the
training pairs (right-hand side, solution) of its inner velocity solves on
seeded random-ellipse viscosity fields, the network is trained on them with
the library's BF16-MFMA forward/backward kernels, and a new field is solved
with and without the network's initial guesses.

What is asserted: training lowers the loss; the learned solve reaches the
same tolerance and the same solution as the plain one (the multigrid CG
finishes every inner solve, so the network can only change the path); the
inner iteration count is reported as the gain (ratio plain / learned) and
may not exceed the plain count by more than 5%.
"""
import numpy as np
import pytest

import mdgpu
from mdgpu import learned as L

pytestmark = pytest.mark.gpu


def test_learned_guess_trains_and_keeps_the_solution():
    n = 63
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(*L.free_slip_bc(n + 1))
    probs = [L.EllipseProblem(ctx, n, seed, n_ell=6) for seed in range(1, 9)]
    stats, npairs = L.collect_pairs(ctx, probs[:7], 2000, rel_tol=1e-8)
    assert npairs == sum(s.its for s in stats)
    losses = L.train(ctx, 1500, batch=8, lr=1e-3, seed=3)
    first, last = losses[:20].mean(), losses[-50:].mean()
    print("pairs", npairs, "loss", first, "->", last)
    assert last < 0.5 * first
    test = probs[7]
    xp, op = test.solve(ctx, precond=0, rel_tol=1e-8)
    xl, ol = test.solve(ctx, precond=1, rel_tol=1e-8)
    print("plain: outer %d inner %d %.1f ms | learned: outer %d inner %d %.1f ms, guesses %d, "
          "mean ||r0||^2/||b||^2 %.3f | inner gain %.3f" %
          (op.its, op.inner_its, op.time_ms, ol.its, ol.inner_its, ol.time_ms, ol.guess_used,
           ol.guess_gain, op.inner_its / max(ol.inner_its, 1)))
    assert op.rel_res <= 1e-8 and ol.rel_res <= 1e-8
    assert ol.guess_used > 0
    assert np.linalg.norm(xl - xp) <= 1e-6 * np.linalg.norm(xp)
    assert ol.inner_its <= 1.05 * op.inner_its


def test_learned_outer_directions_train_and_keep_the_solution():
    """Outer variant: the network is trained on the (residual, error) pairs
    of the plain solver's own outer iterations, first with the error (L2)
    loss and then with the loss the solver sees (the residual left after the
    best combination of its velocity and pressure directions), and is then
    used in both outer integrations: precond = 2 (two more GCR directions per
    iteration) and precond = 3 (network correction first, multigrid on what
    it leaves).  GCR weights every direction optimally, so the learned solves
    reach the same solution; iteration counts and times are reported, and a
    learned solve may not need more than one extra outer iteration."""
    n = 63
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(*L.free_slip_bc(n + 1))
    probs = [L.EllipseProblem(ctx, n, seed, n_ell=6) for seed in range(1, 10)]
    stats, npairs = L.collect_outer_pairs(ctx, probs[:8], 2000, rel_tol=1e-8)
    assert npairs == sum(s.its for s in stats)
    losses = L.train(ctx, 800, batch=8, lr=1e-3, seed=3)
    first, last = losses[:20].mean(), losses[-50:].mean()
    print("outer pairs", npairs, "error loss", first, "->", last)
    assert last < 0.7 * first
    rl = L.train_residual(ctx, 1500, batch=8, lr=1e-3, seed=3)
    r_first, r_last = rl[:20].mean(), rl[-50:].mean()
    print("residual loss", r_first, "->", r_last)
    assert r_last < 0.97 * r_first
    test = probs[8]
    xp, op = test.solve(ctx, precond=0, rel_tol=1e-8)
    print("plain: outer %d inner %d %.1f ms" % (op.its, op.inner_its, op.time_ms))
    assert op.rel_res <= 1e-8
    for pc in (2, 3):
        xl, ol = test.solve(ctx, precond=pc, rel_tol=1e-8)
        print("learned precond %d: outer %d inner %d %.1f ms, network directions %d, residual ratio "
              "of the network part %.3f | outer gain %.3f, time gain %.3f" %
              (pc, ol.its, ol.inner_its, ol.time_ms, ol.guess_used, ol.guess_gain,
               op.its / max(ol.its, 1), op.time_ms / ol.time_ms))
        assert ol.rel_res <= 1e-8
        assert ol.guess_used > 0
        assert np.linalg.norm(xl - xp) <= 1e-6 * np.linalg.norm(xp)
        assert ol.its <= op.its + 1


def test_learned_precond_needs_parameters():
    """precond 1 or 2 without network parameters is an error, not a silent
    fall back to the plain method."""
    n = 31
    ctx = mdgpu.Context(n + 1, n + 1, 0.0, 1.0, 0.0, 1.0)
    ctx.set_bc(*L.free_slip_bc(n + 1))
    p = L.EllipseProblem(ctx, n, 1, n_ell=3)
    for pc in (1, 2):
        with pytest.raises(mdgpu.MdgError) as e:
            p.solve(ctx, precond=pc, rel_tol=1e-8)
        assert e.value.code == -4
