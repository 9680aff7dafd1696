"""GPU parity of the marker index and P2Mastah against the oracle.

Marker -> node indices must be bit-exact; node fields agree to 1e-12
relative (summation order differs from the reference's marker order).
"""
import numpy as np
import pytest

import mdgpu
from mdgpu import setups as S

pytestmark = pytest.mark.gpu


def _case(oracle_mod, n=33, m=27, nper=4, seed=11, box=(-0.5, 0.5, -0.3, 0.45)):
    g = oracle_mod.Grid(n, m, *box)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, nper, nper, seed=seed)
    phase = ((np.sin(9 * x) + np.cos(7 * z)) > 0.3).astype(np.int32)
    phase[::97] = -1                       # inactive markers are skipped
    ctx = mdgpu.Context(n, m, *box)
    ctx.markers_set(x, z, phase)
    return g, ctx, x, z, phase


def test_coordinates_match_oracle(oracle_mod):
    g, ctx, *_ = _case(oracle_mod)
    c = ctx.coords()
    for k in ("xg", "zg", "xc", "zc", "xvz", "zvx"):
        assert np.array_equal(c[k], getattr(g, k)), k


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
def test_marker_index_bit_exact(oracle_mod, centroid):
    g, ctx, x, z, ph = _case(oracle_mod)
    # add markers outside the box and exactly on node midpoints / cell edges
    print("marker index centroid", centroid, flush=True)
    ip, jp = ctx.marker_index(centroid)
    rip, rjp = oracle_mod.marker_index(g, centroid, x, z)
    assert np.array_equal(ip, rip) and np.array_equal(jp, rjp)


def test_marker_index_edges_and_ties(oracle_mod):
    box = (0.0, 8.0, 0.0, 6.0)
    g = oracle_mod.Grid(9, 7, *box)
    xs = np.concatenate([np.arange(-1, 18) * 0.5, [8.0, 0.0, 1e-300, 7.999999999999999]])
    zs = np.linspace(-0.5, 6.5, xs.size)
    ctx = mdgpu.Context(9, 7, *box)
    ctx.markers_set(xs, zs, np.zeros(xs.size, np.int32))
    for centroid in (1, 0, -1, -2):
        ip, jp = ctx.marker_index(centroid)
        rip, rjp = oracle_mod.marker_index(g, centroid, xs, zs)
        assert np.array_equal(ip, rip) and np.array_equal(jp, rjp), centroid


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
@pytest.mark.parametrize("stencil", [1, 9])
@pytest.mark.parametrize("avg", [0, 1, 2])
def test_p2mastah_table_field(oracle_mod, centroid, stencil, avg):
    g, ctx, x, z, ph = _case(oracle_mod)
    print("p2g table", centroid, stencil, avg, flush=True)
    table = np.array([1.0, 1e3])
    nn = g.node_count(centroid)
    ref = oracle_mod.p2mastah(g, x, z, ph, table, flag=0, avg=avg, centroid=centroid,
                              stencil=stencil, node_field=np.full(nn, -7.0))
    got = ctx.p2mastah(table, node_field=np.full(nn, -7.0), flag=0, avg=avg, centroid=centroid,
                       interp_stencil=stencil)
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
@pytest.mark.parametrize("stencil", [1, 9])
def test_p2mastah_per_marker(oracle_mod, centroid, stencil):
    g, ctx, x, z, ph = _case(oracle_mod, nper=3)
    f = np.exp(np.sin(5 * x) * 4 + z)
    nn = g.node_count(centroid)
    for avg in (0, 1, 2):
        ref = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, avg=avg, centroid=centroid,
                                  stencil=stencil)
        got = ctx.p2mastah(f, flag=1, avg=avg, centroid=centroid, interp_stencil=stencil)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)


@pytest.mark.parametrize("centroid", [1, 0])
@pytest.mark.parametrize("stencil", [1, 9])
def test_phase_fractions(oracle_mod, centroid, stencil):
    g, ctx, x, z, ph = _case(oracle_mod)
    ref = oracle_mod.p2mastah(g, x, z, ph, np.zeros(2), prop=1, centroid=centroid,
                              stencil=stencil, nb_phases=2)
    got = ctx.p2mastah(np.zeros(2), prop=1, centroid=centroid, interp_stencil=stencil,
                       nb_phases=2)
    for p in range(2):
        np.testing.assert_allclose(got[p], ref[p], rtol=1e-12, atol=1e-15)


def test_fused_batch_and_node_tags(oracle_mod):
    g, ctx, x, z, ph = _case(oracle_mod)
    nn = g.node_count(1)
    tags = np.full(nn, -1, np.int8)
    tags[5:40] = 30                            # air nodes keep their value
    fields, refs = [], []
    for q in range(6):
        f = 1.0 + (q + 1) * (x - z) ** 2
        dst = np.full(nn, -3.0)
        fields.append(dict(src=f, dst=dst, flag=1, avg=q % 3, node_type=tags))
        refs.append(oracle_mod.p2mastah(g, x, z, ph, f, flag=1, avg=q % 3, centroid=1,
                                        node_type=tags, node_field=np.full(nn, -3.0)))
    ctx.p2mastah_fused(1, 1, fields)
    for f, r in zip(fields, refs):
        np.testing.assert_allclose(f["dst"], r, rtol=1e-12)
        assert np.all(f["dst"][5:40] == -3.0)


def test_device_resident_sorted_fields(oracle_mod):
    g, ctx, x, z, ph = _case(oracle_mod, n=65, m=49)
    xs, zs, ps = ctx.device(len(x), host=x), ctx.device(len(z), host=z), ctx.device(len(ph), np.int32, ph)
    ctx.markers_set(xs, zs, ps)
    f = np.cos(3 * x) + 2.0
    fs = ctx.device(len(x))
    ctx.markers_gather(f, fs)
    out = ctx.device(g.node_count(0), host=np.zeros(g.node_count(0)))
    ctx.p2mastah_fused(0, 1, [dict(src=fs, dst=out, flag=1 | mdgpu.MDG_SRC_SORTED, avg=0)])
    ref = oracle_mod.p2mastah(g, x, z, ph, f, flag=1, centroid=0)
    np.testing.assert_allclose(out.numpy(), ref, rtol=1e-12)
    # a second call on the same resident data is bit-identical (no atomics)
    out2 = ctx.device(g.node_count(0), host=np.zeros(g.node_count(0)))
    ctx.p2mastah_fused(0, 1, [dict(src=fs, dst=out2, flag=1 | mdgpu.MDG_SRC_SORTED, avg=0)])
    assert np.array_equal(out.numpy(), out2.numpy())


@pytest.mark.parametrize("stencil", [1, 9])
def test_p2mastah_on_ties(oracle_mod, stencil):
    """Markers exactly on cell edges, node midpoints and the box walls, where
    a marker's nearest node can differ from the one its sort key implies
    (the node set then takes the path that recomputes every marker's node):
    P2Mastah matches the oracle on every node set with both stencils, per
    marker and through a phase table."""
    box = (0.0, 8.0, 0.0, 6.0)
    g = oracle_mod.Grid(9, 7, *box)
    xs = np.repeat(np.arange(0, 17) * 0.5, 13)
    zs = np.tile(np.arange(0, 13) * 0.5, 17)
    ph = (np.arange(xs.size) % 2).astype(np.int32)
    f = 1.5 + np.sin(xs) * np.cos(zs)
    ctx = mdgpu.Context(9, 7, *box)
    ctx.markers_set(xs, zs, ph)
    for centroid in (1, 0, -1, -2):
        for avg in (0, 2):
            ref = oracle_mod.p2mastah(g, xs, zs, ph, f, flag=1, avg=avg, centroid=centroid, stencil=stencil)
            got = ctx.p2mastah(f, flag=1, avg=avg, centroid=centroid, interp_stencil=stencil)
            np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0, err_msg="centroid %d" % centroid)
        table = np.array([2.0, 5.0])
        ref = oracle_mod.p2mastah(g, xs, zs, ph, table, flag=0, avg=1, centroid=centroid, stencil=stencil)
        got = ctx.p2mastah(table, flag=0, avg=1, centroid=centroid, interp_stencil=stencil)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0, err_msg="table, centroid %d" % centroid)


@pytest.mark.parametrize("centroid", [1, 0, -1, -2])
@pytest.mark.parametrize("stencil", [1, 9])
def test_cached_weights_follow_new_markers(oracle_mod, centroid, stencil):
    """The per-marker state cached by the first P2G call of a node set (the
    1-cell weights and whether every marker sits in the home its sort key
    names, which the 9-cell stencil's fast path relies on) is recomputed
    after mdg_markers_set: P2G on a second marker set (moved markers, same
    context) matches the oracle on that set, and a repeated call on it is
    bit-identical."""
    g, ctx, x, z, phase = _case(oracle_mod)
    f = np.cos(3 * x) * np.sin(2 * z) + 2.0
    nn = g.node_count(centroid)
    ctx.p2mastah(f, flag=1, avg=0, centroid=centroid, interp_stencil=stencil)
    rng = np.random.default_rng(3)
    x2 = x + 0.3 * g.dx * rng.standard_normal(x.size)
    z2 = z + 0.3 * g.dz * rng.standard_normal(z.size)
    ctx.markers_set(x2, z2, phase)
    f2 = np.sin(5 * x2) + 3.0
    for avg in (0, 2):
        ref = oracle_mod.p2mastah(g, x2, z2, phase, f2, flag=1, avg=avg, centroid=centroid,
                                  stencil=stencil)
        got = ctx.p2mastah(f2, flag=1, avg=avg, centroid=centroid, interp_stencil=stencil)
        again = ctx.p2mastah(f2, flag=1, avg=avg, centroid=centroid, interp_stencil=stencil)
        assert got.shape == (nn,)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
        assert np.array_equal(got, again)
