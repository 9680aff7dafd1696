"""Pins of the oracle's grid, marker index and P2Mastah restatement.

Each check is independent of the oracle's code path: brute-force nearest
nodes, a pure-Python re-derivation of the distance-weighted averages on tiny
grids, closed forms (constant fields, partition of unity, harmonic and
geometric means), and the reference's own InterpEquiv property (thread
buffers vs atomic scatter give the same fields, TESTS/InterpEquivalenceTests.cpp).
"""
import itertools

import numpy as np
import pytest

from mdgpu import setups as S


def _grid(oracle_mod, n=9, m=7, box=(-0.5, 0.5, -0.3, 0.4)):
    return oracle_mod.Grid(n, m, *box)


def test_coordinates_are_running_sums(oracle_mod):
    g = oracle_mod.Grid(103, 57, -0.5, 0.5, -0.25, 0.3)
    ref = S.grid_coords(103, 57, -0.5, 0.5, -0.25, 0.3)
    for k in ("xg", "zg", "xc", "zc", "xvz", "zvx"):
        assert np.array_equal(ref[k], getattr(g, k)), k
    # a running sum is not xmin + i*dx: the two differ in the last bits
    assert not np.array_equal(g.xg, -0.5 + np.arange(103) * g.dx)


def _brute_nearest(X, pos):
    d = np.abs(X[None, :] - pos[:, None])
    return np.argmin(d, axis=1)   # first minimum = lower index on ties


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
def test_marker_index_matches_nearest_node(oracle_mod, centroid):
    g = _grid(oracle_mod, 17, 13)
    rng = np.random.default_rng(3)
    x = rng.uniform(-0.6, 0.6, 5000)
    z = rng.uniform(-0.4, 0.5, 5000)
    ip, jp = oracle_mod.marker_index(g, centroid, x, z)
    X, Z = {1: (g.xc, g.zc), 0: (g.xg, g.zg), -1: (g.xg, g.zvx), -2: (g.xvz, g.zg)}[centroid]
    assert np.array_equal(ip, _brute_nearest(X, x))
    assert np.array_equal(jp, _brute_nearest(Z, z))


def test_marker_index_ties_go_to_lower_node(oracle_mod):
    g = oracle_mod.Grid(5, 5, 0.0, 4.0, 0.0, 4.0)     # dx = 1: midpoints exact
    x = np.array([0.5, 1.5, 2.5, 3.5, -3.0, 9.0])
    z = np.array([1.5, 0.5, 3.5, 2.5, 0.0, 4.0])
    ip, jp = oracle_mod.marker_index(g, 0, x, z)
    assert list(ip) == [0, 1, 2, 3, 0, 4]           # ties down, clamped at both ends
    assert list(jp) == [1, 0, 3, 2, 0, 4]


def _py_p2g(g, x, z, phase, vals, centroid, stencil, avg, flag):
    """Direct transcription of the weight definition (ParticleRoutines.c:2218-2221,
    2357-2576) with explicit loops: nearest node, then the node and its
    in-range neighbours, weight (1-|dx|/dxi)(1-|dz|/dzi)."""
    X, Z = {1: (g.xc, g.zc), 0: (g.xg, g.zg), -1: (g.xg, g.zvx), -2: (g.xvz, g.zg)}[centroid]
    nx, nz = len(X), len(Z)
    dxi = g.dx / 2 if stencil == 1 else 1.5 * g.dx
    dzi = g.dz / 2 if stencil == 1 else 1.5 * g.dz
    W = np.zeros(nx * nz); BW = np.zeros(nx * nz)
    for k in range(len(x)):
        ip = int(_brute_nearest(X, np.array([x[k]]))[0])
        jp = int(_brute_nearest(Z, np.array([z[k]]))[0])
        v = vals[phase[k]] if flag == 0 else vals[k]
        if avg == 1:
            v = 1.0 / v
        elif avg == 2:
            v = np.log(v)
        offs = [(0, 0)] if stencil == 1 else list(itertools.product((-1, 0, 1), repeat=2))
        for ox, oz in offs:
            i, j = ip + ox, jp + oz
            if not (0 <= i < nx and 0 <= j < nz):
                continue
            w = (1 - abs(X[ip] + ox * g.dx - x[k]) / dxi) * (1 - abs(Z[jp] + oz * g.dz - z[k]) / dzi)
            W[i + j * nx] += w
            BW[i + j * nx] += v * w
    out = np.full(nx * nz, np.nan)
    m = W >= 1e-30
    out[m] = BW[m] / W[m]
    if avg == 1:
        out = 1.0 / out
    if avg == 2:
        out = np.exp(out)
    return out, W


def _markers(g, nper=3, seed=5):
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, nper, nper, seed=seed)
    phase = ((x * 3.1 + z * 1.7) > 0.05).astype(np.int32) + (x > 0.2).astype(np.int32)
    return x, z, phase


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
@pytest.mark.parametrize("stencil", [1, 9])
@pytest.mark.parametrize("avg", [0, 1, 2])
def test_p2mastah_matches_bruteforce(oracle_mod, centroid, stencil, avg):
    g = _grid(oracle_mod)
    x, z, ph = _markers(g)
    table = np.array([1.0, 7.5, 0.125])
    ref, W = _py_p2g(g, x, z, ph, table, centroid, stencil, avg, 0)
    nf = np.full(g.node_count(centroid), -99.0)
    got = oracle_mod.p2mastah(g, x, z, ph, table, flag=0, avg=avg, centroid=centroid,
                              stencil=stencil, node_field=nf)
    m = W >= 1e-30
    np.testing.assert_allclose(got[m], ref[m], rtol=1e-13)
    assert np.all(got[~m] == -99.0)      # empty nodes keep their previous value


def test_p2mastah_per_marker_field_and_identities(oracle_mod):
    g = _grid(oracle_mod, 12, 10)
    x, z, ph = _markers(g, 4)
    f = np.exp(np.sin(7 * x) + np.cos(5 * z))
    ari = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, avg=0, centroid=0, stencil=9)
    har = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, avg=1, centroid=0, stencil=9)
    geo = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, avg=2, centroid=0, stencil=9)
    assert np.all(har <= geo * (1 + 1e-14)) and np.all(geo <= ari * (1 + 1e-14))   # HM <= GM <= AM
    c = oracle_mod.p2mastah(g, x, z, ph, np.full(len(x), 3.25), flag=1, avg=2, centroid=1)
    np.testing.assert_allclose(c, 3.25, rtol=1e-14)


@pytest.mark.parametrize("centroid", [0, 1])
def test_phase_fractions_partition_of_unity(oracle_mod, centroid):
    g = _grid(oracle_mod, 10, 8)
    x, z, ph = _markers(g)
    perc = oracle_mod.p2mastah(g, x, z, ph, np.zeros(3), prop=1, centroid=centroid, nb_phases=3)
    np.testing.assert_allclose(perc.sum(axis=0), 1.0, rtol=1e-14)
    # brute force: fraction of weight carried by each phase
    for p in range(3):
        ind = (ph == p).astype(float)
        ref, W = _py_p2g(g, x, z, np.zeros(len(x), np.int32), ind, centroid, 1, 0, 1)
        np.testing.assert_allclose(perc[p], ref, rtol=1e-12, atol=1e-15)


@pytest.mark.parametrize("stencil", [1, 9])
def test_interp_modes_agree(oracle_mod, stencil):
    """InterpEquivalenceTests.cpp: thread-private buffers (mode 0) and the fused
    atomic scatter (mode 3) give the same node fields."""
    g = _grid(oracle_mod, 40, 30)
    x, z, ph = _markers(g, 4)
    f = 1.0 + x * x
    a = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, centroid=1, stencil=stencil, mode=0, nthreads=4)
    b = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, centroid=1, stencil=stencil, mode=3, nthreads=4)
    np.testing.assert_allclose(a, b, rtol=1e-13)
    s = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, centroid=1, stencil=stencil, mode=0, nthreads=1)
    np.testing.assert_allclose(a, s, rtol=1e-13)
