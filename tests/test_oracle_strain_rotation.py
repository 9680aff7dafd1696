"""Oracle pins of the marker strain accumulation (AccumulatedStrainII,
RheologyParticles.c:346-649) and of the marker stress rotation
(UpdateParticleStress, :1965-1988), from what the mathematics fixes:

* the increments are interpolated with weights that sum to one (left-out air
  centres are renormalised away), so constant rates add exactly dt times
  the rate (to round-off) to every active marker, and a bilinear rate field
  a + b x + c z + d x z is reproduced at every active marker, the clamped
  edge cells extrapolating it exactly;
* the total is the sum of the seven mechanisms, a negative total is refused;
* the analytical rotation (mode 2) is R T R^T with R the rotation by
  dt*w: it keeps the trace and the second invariant, and rotating back by
  -dt*w restores the tensor; the Jaumann rate (mode 1) agrees to second
  order in dt*w with the rotation by -dt*w (the reference's two modes turn
  the tensor in opposite senses)."""
import numpy as np
import pytest

from mdgpu import setups as S


def _grid(O, n=21, seed=2):
    g = O.Grid(n, n + 4, 0.0, 1.0, -1.2, 0.0)
    g.pure_shear(1.0)
    x, z = S.put_markers(g.xg, g.zg, g.dx, g.dz, 3, 3, seed=seed)
    ph = np.zeros(x.size, np.int32)
    ph[::17] = -1
    return g, x, z, ph


def test_constant_and_bilinear_rates(oracle_mod):
    O = oracle_mod
    g, x, z, ph = _grid(O)
    nc = (g.Nx - 1) * (g.Nz - 1)
    XC, ZC = np.meshgrid(g.xc, g.zc)
    a, b, c, d = 0.3, 1.7, -0.4, 0.9
    lin = (a + b * XC + c * ZC + d * XC * ZC).ravel()
    dt = 0.25
    rates = {"eII_pl": np.full(nc, 2.0), "eII_el": lin, "eII_pwl": np.full(nc, 0.5)}
    zero = {k: np.zeros(x.size) for k in O.STRAINS}
    out = O.accumulated_strain(g, dt, x, z, ph, rates, zero, div_u_pl=lin)
    on = ph != -1
    np.testing.assert_allclose(out["strain_pl"][on], dt * 2.0, rtol=1e-14)
    np.testing.assert_allclose(out["strain_pwl"][on], dt * 0.5, rtol=1e-14)
    exact = a + b * x + c * z + d * x * z
    np.testing.assert_allclose(out["strain_el"][on], dt * exact[on], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(out["strain_pl_vol"][on], dt * exact[on], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(out["strain"][on], dt * (2.0 + 0.5 + exact[on]), rtol=1e-12, atol=1e-14)
    for k in ("strain_exp", "strain_lin", "strain_gbs"):
        assert np.all(out[k] == 0.0)
    assert np.all(out["strain"][~on] == 0.0)
    # an air row (BCp 30) is left out and the weights renormalised: a
    # constant is still reproduced next to it, wherever the remaining weight
    # is positive (above the air row's centres; below them the clamped cell
    # leaves a negative weight sum, which the reference does not normalise)
    g.BCp_t[: g.Nx - 1] = 30
    out = O.accumulated_strain(g, dt, x, z, ph, {"eII_pl": np.full(nc, 2.0)}, {"strain_pl": np.zeros(x.size)})
    above = on & (z > g.zc[0] + 1e-9)
    np.testing.assert_allclose(out["strain_pl"][above], dt * 2.0, rtol=1e-14)
    with pytest.raises(ValueError):
        O.accumulated_strain(g, dt, x, z, ph, {"eII_pl": np.full(nc, -1.0)}, {"strain": np.zeros(x.size)})


def test_rotation_invariants_and_jaumann(oracle_mod):
    O = oracle_mod
    rng = np.random.default_rng(4)
    n = 1000
    ph = np.zeros(n, np.int32)
    ph[::9] = -1
    txx, tzz, txz = rng.standard_normal((3, n))
    w = rng.standard_normal(n)
    dt = 0.3
    rx, rz, rxz = O.rotate_stress(2, dt, ph, w, txx, tzz, txz)
    on = ph != -1
    np.testing.assert_allclose((rx + rz)[on], (txx + tzz)[on], atol=1e-14)
    np.testing.assert_allclose((rx ** 2 + rz ** 2 + 2 * rxz ** 2)[on], (txx ** 2 + tzz ** 2 + 2 * txz ** 2)[on],
                               rtol=1e-13)
    bx, bz, bxz = O.rotate_stress(2, -dt, ph, w, rx, rz, rxz)
    np.testing.assert_allclose(np.stack([bx, bz, bxz])[:, on], np.stack([txx, tzz, txz])[:, on], atol=1e-14)
    assert np.array_equal(rx[~on], txx[~on])
    # R T R^T against the same product written with matrices
    k = np.flatnonzero(on)[3]
    t = dt * w[k]
    R = np.array([[np.cos(t), np.sin(t)], [-np.sin(t), np.cos(t)]])
    T = np.array([[txx[k], txz[k]], [txz[k], tzz[k]]])
    M = R @ T @ R.T
    np.testing.assert_allclose([rx[k], rz[k], rxz[k]], [M[0, 0], M[1, 1], M[0, 1]], rtol=1e-13, atol=1e-15)
    # the Jaumann rate is the first order of the rotation by -dt*w: the two
    # modes of the reference turn the tensor in opposite senses
    for eps in (1e-3, 1e-4):
        jx, jz, jxz = O.rotate_stress(1, eps, ph, w, txx, tzz, txz)
        ax, az, axz = O.rotate_stress(2, -eps, ph, w, txx, tzz, txz)
        err = np.abs(np.stack([jx - ax, jz - az, jxz - axz])[:, on]).max()
        assert err < 4 * (eps * np.abs(w).max()) ** 2 * np.abs(np.stack([txx, tzz, txz])).max()
