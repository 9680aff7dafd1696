"""Oracle pin of the marker re-seeding (CountPartCell_OLD,
ParticleRoutines.c:3224-4050) against an independent brute force: fine cells
of dx/2 x dz/2 on the box plus one coarse cell of margin, markers counted by
floor of their offset (no marker sits on a fine-cell edge in these cases, so
this equals the reference's nearest-centre rule), every interior non-air fine
cell with fewer than min_part_cell / 4 markers gets one marker at its centre
that copies the nearest active marker of the 5 x 5 fine cells around it, new
markers in x-major order take the inactive slots first, then are appended.
Error paths: an empty neighbourhood (the reference exits) and a full marker
array."""
import numpy as np
import pytest

from mdgpu import setups as S


# (dk, dl) of the 25 fine cells in the order the reference lists their markers
WINDOW = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1), (-1, 1), (1, 1), (-2, 0), (2, 0), (0, -2),
          (0, 2), (-2, -2), (2, -2), (-2, 2), (2, 2), (-1, -2), (1, -2), (-1, 2), (1, 2), (-2, -1), (-2, 1), (2, -1),
          (2, 1)]


def _brute(g, x, z, ph, min_part_cell):
    Ncx, Ncz = g.Nx - 1, g.Nz - 1
    nfx, nfz = 2 * (Ncx + 2), 2 * (Ncz + 2)
    x0, z0 = g.xg[0] - g.dx, g.zg[0] - g.dz
    act = np.nonzero(ph != -1)[0]
    fi = np.floor((x[act] - x0) / (g.dx / 2)).astype(int).clip(0, nfx - 1)
    fj = np.floor((z[act] - z0) / (g.dz / 2)).astype(int).clip(0, nfz - 1)
    cnt = np.bincount(fi + fj * nfx, minlength=nfx * nfz)
    thr = min_part_cell // 4
    out = []
    for k in range(2, nfx - 2):
        for l in range(2, nfz - 2):
            if g.BCp_t[(k - 2) // 2 + ((l - 2) // 2) * Ncx] == 30 or cnt[k + l * nfx] >= thr:
                continue
            sel = (np.abs(fi - k) <= 2) & (np.abs(fj - l) <= 2)
            if not sel.any():
                return None
            nx_ = g.xg[0] + (k - 2) * g.dx / 2.0 + g.dx / 4.0
            nz_ = g.zg[0] + (l - 2) * g.dz / 2.0 + g.dz / 4.0
            cand = act[sel]
            ddx, ddz = nx_ - x[cand], nz_ - z[cand]
            d = np.sqrt(ddx * ddx + ddz * ddz)
            # ties: the first in the reference's list order (window rank, then index)
            rank = np.array([WINDOW.index((a, b)) for a, b in zip(fi[sel] - k, fj[sel] - l)])
            best = np.lexsort((cand, rank, d))[0]
            out.append((nx_, nz_, cand[best]))
    return out


def _case(O, n=24, seed=3, hole=(0.3, 0.37, -0.7, -0.63), jitter=0.3, inactive=0.02, drop=False):
    g = O.Grid(n + 1, n + 1, 0.0, 1.0, -1.0, 0.0)
    rng = np.random.default_rng(seed)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=seed)
    x = x + rng.uniform(-jitter, jitter, x.size) * g.dx / 4
    z = z + rng.uniform(-jitter, jitter, z.size) * g.dz / 4
    ph = rng.integers(0, 3, x.size).astype(np.int32)
    inside = (x > hole[0]) & (x < hole[1]) & (z > hole[2]) & (z < hole[3])
    ph[inside] = -1
    ph[rng.random(x.size) < inactive] = -1
    if drop:                       # the hole's markers gone from the arrays, not tagged inactive
        x, z, ph = x[~inside], z[~inside], ph[~inside]
    return g, x, z, ph


@pytest.mark.parametrize("seed,mpc", [(3, 16), (4, 16), (5, 20), (6, 12)])
def test_reseed_matches_brute_force(oracle_mod, seed, mpc):
    O = oracle_mod
    g, x, z, ph = _case(O, seed=seed)
    expect = _brute(g, x, z, ph, mpc)
    X, Z, P, slot, src, rc = O.reseed(g, x, z, ph, min_part_cell=mpc)
    assert rc == len(expect) > 10
    inactive = np.nonzero(ph == -1)[0]
    k = min(rc, inactive.size)
    assert np.array_equal(slot[:k], inactive[:k])
    assert np.array_equal(slot[k:], x.size + np.arange(rc - k))
    assert X.size == x.size + (rc - k)
    for i, (ex, ez, es) in enumerate(expect):
        assert src[i] == es
        assert X[slot[i]] == ex and Z[slot[i]] == ez and P[slot[i]] == ph[es]
    untouched = np.setdiff1d(np.arange(x.size), slot)
    assert np.array_equal(X[untouched], x[untouched]) and np.array_equal(P[untouched], ph[untouched])


def test_equidistant_markers_follow_the_list_order(oracle_mod):
    """Markers on the regular 4 x 4 lattice: a new marker at a fine-cell
    centre is equidistant from several markers, and the reference keeps the
    first of them in its list order."""
    O = oracle_mod
    g = O.Grid(17, 17, 0.0, 1.0, -1.0, 0.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 4, 4, seed=0, noise=False)
    ph = (np.arange(x.size) % 5).astype(np.int32)
    ph[(x > 0.4) & (x < 0.5) & (z > -0.6) & (z < -0.5)] = -1
    expect = _brute(g, x, z, ph, 16)
    X, Z, P, slot, src, rc = O.reseed(g, x, z, ph)
    assert rc == len(expect) > 0
    assert [int(s_) for s_ in src] == [int(e[2]) for e in expect]


def test_air_cells_are_not_reseeded(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _case(O, hole=(0.0, 1.0, -0.12, 0.0))  # empty top band
    Ncx = g.Nx - 1
    g.BCp_t[(g.Nz - 2) * Ncx:] = 30                        # top row of cells is air
    g.BCp_t[(g.Nz - 3) * Ncx:] = 30                        # and the one below
    expect = _brute(g, x, z, ph, 16)
    X, Z, P, slot, src, rc = O.reseed(g, x, z, ph)
    assert rc == len(expect)
    assert np.all(Z[slot] < g.zg[-3])


def test_error_paths(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _case(O, hole=(0.2, 0.8, -0.8, -0.2))   # a hole wider than the 5 x 5 window
    assert _brute(g, x, z, ph, 16) is None
    assert O.reseed(g, x, z, ph)[-1] == -1
    g, x, z, ph = _case(O, inactive=0.002, drop=True)
    n_new = len(_brute(g, x, z, ph, 16))
    n_free = int(np.sum(ph == -1))
    assert n_new > n_free
    assert O.reseed(g, x, z, ph, capacity=x.size + n_new - n_free - 1)[-1] == -2
    assert O.reseed(g, x, z, ph, capacity=x.size + n_new - n_free)[-1] == n_new
