"""GPU marker re-seeding (mdg_count_part_cell) against the oracle's
restatement of CountPartCell_OLD (ParticleRoutines.c:3224-4050): integer
and index work, compared bit for bit (new marker count, slots, source
markers, coordinates and phases of every marker).  Cases: holes and inactive
markers on jittered layouts at three thresholds, the regular lattice where
new markers are equidistant from several markers (the reference's list
order decides), air cells, device-resident markers, the two error paths, and
the 1023^2 grid with 16 markers per cell."""
import time

import numpy as np
import pytest

import mdgpu
from test_oracle_reseed import _case

pytestmark = pytest.mark.gpu


def _ctx(g):
    # the box the oracle grids of these cases are built on (its dx, dz)
    return mdgpu.Context(g.Nx, g.Nz, 0.0, 1.0, -1.0, 0.0)


def _padded(x, z, ph, cap):
    X, Z, P = np.zeros(cap), np.zeros(cap), np.full(cap, -1, np.int32)
    X[:x.size], Z[:x.size], P[:x.size] = x, z, ph
    return X, Z, P


def _compare(O, g, x, z, ph, mpc=16, device=False, ctx=None):
    cap = x.size + 4 * g.Nx * g.Nz
    Xo, Zo, Po, slot_o, src_o, rc = O.reseed(g, x, z, ph, min_part_cell=mpc, capacity=cap)
    assert rc >= 0
    ctx = ctx or _ctx(g)
    X, Z, P = _padded(x, z, ph, cap)
    t = time.time()
    if device:
        dX, dZ, dP = (mdgpu.DeviceArray(ctx, a.size, a.dtype, host=a) for a in (X, Z, P))
        n_out, slot, src = ctx.count_part_cell(dX, dZ, dP, n=x.size, min_part_cell=mpc, cell_tag=g.BCp_t)
        X, Z, P = dX.numpy(), dZ.numpy(), dP.numpy()
    else:
        n_out, slot, src = ctx.count_part_cell(X, Z, P, n=x.size, min_part_cell=mpc, cell_tag=g.BCp_t)
    wall = time.time() - t
    assert n_out == Xo.size
    assert np.array_equal(slot, slot_o) and np.array_equal(src, src_o)
    assert np.array_equal(X[:n_out], Xo) and np.array_equal(Z[:n_out], Zo) and np.array_equal(P[:n_out], Po)
    return rc, ctx, wall


@pytest.mark.parametrize("seed,mpc", [(3, 16), (4, 16), (5, 20), (6, 12)])
def test_reseed_matches_oracle(oracle_mod, seed, mpc):
    g, x, z, ph = _case(oracle_mod, seed=seed)
    rc, _, _ = _compare(oracle_mod, g, x, z, ph, mpc)
    assert rc > 10


def test_reseed_device_arrays_and_appends(oracle_mod):
    g, x, z, ph = _case(oracle_mod, inactive=0.002, drop=True)
    rc, _, _ = _compare(oracle_mod, g, x, z, ph, device=True)
    assert rc > int(np.sum(ph == -1))          # some markers appended


def test_reseed_ties_follow_the_list_order(oracle_mod):
    from mdgpu import setups as S
    O = oracle_mod
    g = O.Grid(17, 17, 0.0, 1.0, -1.0, 0.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=0, noise=False)
    ph = (np.arange(x.size) % 5).astype(np.int32)
    ph[(x > 0.4) & (x < 0.5) & (z > -0.6) & (z < -0.5)] = -1
    rc, _, _ = _compare(O, g, x, z, ph)
    assert rc == 9


def test_reseed_air_and_errors(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _case(O, hole=(0.0, 1.0, -0.12, 0.0))
    Ncx = g.Nx - 1
    g.BCp_t[(g.Nz - 3) * Ncx:] = 30
    _compare(O, g, x, z, ph)
    # empty 5 x 5 neighbourhood: refused, markers untouched
    g, x, z, ph = _case(O, hole=(0.2, 0.8, -0.8, -0.2))
    ctx = _ctx(g)
    X, Z, P = _padded(x, z, ph, x.size + 4 * g.Nx * g.Nz)
    X0, P0 = X.copy(), P.copy()
    with pytest.raises(mdgpu.MdgError) as e:
        ctx.count_part_cell(X, Z, P, n=x.size)
    assert e.value.code == -9 and np.array_equal(X, X0) and np.array_equal(P, P0)
    # capacity one short of what the appended markers need
    g, x, z, ph = _case(O, inactive=0.002, drop=True)
    _, _, _, slot, _, rc = O.reseed(g, x, z, ph)
    need = int(slot.max()) + 1
    ctx = _ctx(g)
    X, Z, P = _padded(x, z, ph, need - 1)
    with pytest.raises(mdgpu.MdgError) as e:
        ctx.count_part_cell(X, Z, P, n=x.size)
    assert e.value.code == -1
    X, Z, P = _padded(x, z, ph, need)
    n_out, slot_g, _ = ctx.count_part_cell(X, Z, P, n=x.size)
    assert n_out == need and np.array_equal(slot_g, slot)


def test_reseed_1023(oracle_mod):
    """The 1023^2 grid with 16 markers per cell (16.7M), 300 holes of two
    cells and 1% inactive markers: bit-exact against the oracle, time printed."""
    from mdgpu import setups as S
    O = oracle_mod
    n = 1023
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=11)
    rng = np.random.default_rng(12)
    ph = rng.integers(0, 3, x.size).astype(np.int32)
    ph[rng.random(x.size) < 0.01] = -1
    ci = np.minimum(np.floor(x / g.dx).astype(np.int64), n - 1) + \
        np.minimum(np.floor(-z / g.dz).astype(np.int64), n - 1) * n
    hole = np.zeros(n * n, bool)
    for a, b in zip(rng.integers(0, n - 1, 300), rng.integers(0, n, 300)):
        hole[b * n + a] = hole[b * n + a + 1] = True
    ph[hole[ci]] = -1
    t = time.time()
    rc, ctx, wall = _compare(O, g, x, z, ph)
    print("1023^2 re-seeding: %d new markers, GPU call %.1f ms (kernels %.1f ms), total with oracle %.1f s"
          % (rc, wall * 1e3, ctx.last_kernel_ms(), time.time() - t))
    assert rc >= 8 * 300
