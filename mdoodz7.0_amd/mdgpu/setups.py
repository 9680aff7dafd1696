"""Synthetic workloads: grids, markers, inclusion/ellipse phases, BCs.

The marker layout follows PutPartInBox (ParticleRoutines.c:1102-1194): per
cell (x-index outer, z-index inner), Nx_part x Nz_part markers at the
sub-cell centres plus a jitter of at most 0.1 of the marker spacing
(ParticleRoutines.c:1170-1180).  The reference draws the jitter with libc
rand(); here it comes from a counter-based generator (splitmix64 of
seed and marker index) so every run and every device sees the same markers.
"""
import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(v):
    """splitmix64 finaliser over a uint64 array (Steele, Lea & Flood 2014)."""
    with np.errstate(over="ignore"):
        z = (v + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed, idx):
    """U[0,1) doubles from (seed, counter)."""
    with np.errstate(over="ignore"):
        h = splitmix64(np.asarray(idx, np.uint64) ^ splitmix64(np.uint64(seed) * np.uint64(0x632BE59BD9B4E019) & _M64))
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def grid_coords(Nx, Nz, xmin, xmax, zmin, zmax):
    """Running-sum coordinates (GridRoutines.c:246-330)."""
    dx = (xmax - xmin) / (Nx - 1)
    dz = (zmax - zmin) / (Nz - 1)

    def run(x0, h, n):
        a = np.empty(n)
        a[0] = x0
        a[1:] = h
        return np.cumsum(a)
    return dict(dx=dx, dz=dz, xg=run(xmin, dx, Nx), zg=run(zmin, dz, Nz),
                xc=run(xmin + dx / 2.0, dx, Nx - 1), zc=run(zmin + dz / 2.0, dz, Nz - 1),
                xvz=run(xmin - dx / 2.0, dx, Nx + 1), zvx=run(zmin - dz / 2.0, dz, Nz + 1))


def put_markers(xg, zg, dx, dz, nxp, nzp, seed=0, noise=True):
    """Marker coordinates in the PutPartInBox order (cell column j outer, row i inner)."""
    ncx, ncz = len(xg) - 1, len(zg) - 1
    dxp, dzp = dx / nxp, dz / nzp
    j = np.arange(ncx).reshape(ncx, 1, 1, 1)
    i = np.arange(ncz).reshape(1, ncz, 1, 1)
    kj = np.arange(nxp).reshape(1, 1, nxp, 1)
    ki = np.arange(nzp).reshape(1, 1, 1, nzp)
    x = (xg[j] + dxp * (kj + 0.5)) + 0.0 * ki + 0.0 * i
    z = (zg[i] + dzp * (ki + 0.5)) + 0.0 * kj + 0.0 * j
    x = x.reshape(-1).copy()
    z = z.reshape(-1).copy()
    if noise:
        m = np.arange(x.size, dtype=np.uint64)
        rx = uniform(seed, 2 * m)
        rz = uniform(seed, 2 * m + np.uint64(1))
        x += (2 * rx - 1) * dxp * 0.1
        z += (2 * rz - 1) * dzp * 0.1
    return x, z


def random_ellipses(seed, n, dx, xmin, xmax, zmin, zmax, axis_cells=(5.0, 60.0),
                    logeta=(-3.0, 3.0), rho=(2700.0, 3300.0)):
    """n ellipses: centre uniform in the box, full axes uniform in axis_cells*dx
    (semi-axes are half of that), angle U[0, pi), log10 eta and density uniform."""
    k = np.arange(n, dtype=np.uint64) * np.uint64(8)
    u = [uniform(seed + 7919, k + np.uint64(q)) for q in range(7)]
    return dict(
        xc=xmin + (xmax - xmin) * u[0], zc=zmin + (zmax - zmin) * u[1],
        a=0.5 * dx * (axis_cells[0] + (axis_cells[1] - axis_cells[0]) * u[2]),
        b=0.5 * dx * (axis_cells[0] + (axis_cells[1] - axis_cells[0]) * u[3]),
        theta=np.pi * u[4], logeta=logeta[0] + (logeta[1] - logeta[0]) * u[5],
        rho=rho[0] + (rho[1] - rho[0]) * u[6])


def ellipse_phase(x, z, ell, chunk=1 << 22):
    """Phase = 1 + index of the last ellipse containing the marker, 0 = matrix."""
    ph = np.zeros(x.size, np.int32)
    for s in range(0, x.size, chunk):
        xs, zs = x[s:s + chunk], z[s:s + chunk]
        p = ph[s:s + chunk]
        for e in range(len(ell["xc"])):
            c, sn = np.cos(ell["theta"][e]), np.sin(ell["theta"][e])
            xr, zr = xs - ell["xc"][e], zs - ell["zc"][e]
            u = (c * xr + sn * zr) / ell["a"][e]
            v = (-sn * xr + c * zr) / ell["b"][e]
            p[(u * u + v * v) <= 1.0] = e + 1
    return ph


def solvi_analytic(x, z, rc=0.2, mm=1.0, mc=1e3, er=-1.0):
    """Schmid & Podladchikov (2003) circular-inclusion solution under pure
    shear (strain rate er), written with complex potentials.  Returns
    vx, vz, p (same shape as x)."""
    Z = np.asarray(x, np.float64) + 1j * np.asarray(z, np.float64)
    A = mm * (mc - mm) / (mc + mm)
    inside = np.abs(Z) <= rc
    vx = np.zeros(Z.shape); vz = np.zeros(Z.shape); p = np.zeros(Z.shape)
    Zi = Z[inside]
    V = (mm / (mc + mm)) * 2.0 * er * np.conj(Zi)
    vx[inside], vz[inside] = V.real, V.imag
    Zo = Z[~inside]
    phi = -2.0 * er * A * rc ** 2 / Zo
    dphi = 2.0 * er * A * rc ** 2 / Zo ** 2
    psi = -2.0 * er * mm * Zo - 2.0 * er * A * rc ** 4 / Zo ** 3
    V = (phi - Zo * np.conj(dphi) - np.conj(psi)) / (2.0 * mm)
    vx[~inside], vz[~inside] = V.real, V.imag
    p[~inside] = (-2.0 * mm * (mc - mm) / (mc + mm) * (rc ** 2 / Zo ** 2 * 2.0 * er)).real
    return vx, vz, p


def weak_seeds(seed, n, radius, xmin, xmax, zmin, zmax):
    """n seeded circle centres (uniform in the box) of the given radius."""
    k = np.arange(n, dtype=np.uint64)
    cx = xmin + (xmax - xmin) * uniform(seed + 101, 2 * k)
    cz = zmin + (zmax - zmin) * uniform(seed + 101, 2 * k + np.uint64(1))
    return np.stack([cx, cz, np.full(n, radius)], axis=1)


def seed_phase(x, z, seeds, chunk=1 << 22):
    """Phase 1 inside any seed circle, 0 elsewhere."""
    ph = np.zeros(x.size, np.int32)
    for s in range(0, x.size, chunk):
        xs, zs = x[s:s + chunk], z[s:s + chunk]
        inside = np.zeros(xs.size, bool)
        for cx, cz, r in seeds:
            inside |= (xs - cx) ** 2 + (zs - cz) ** 2 < r * r
        ph[s:s + chunk][inside] = 1
    return ph


def vep_phases(n_pwl=3.0, C=1.0, C_weak_ratio=0.5, fric_deg=30.0, eta_pwl=1.0):
    """Nondimensional visco-elasto-plastic shear-box phases: a power-law
    matrix (B = eta_pwl at unit strain rate) with Drucker-Prager yield, and
    weak seeds with a reduced cohesion.  Returned as keyword dicts of
    mdg_phase_prop fields."""
    import math
    base = dict(cstv=0, pwlv=1, plast=1, eta0=1.0, pref_pwl=1.0, F_pwl=1.0,
                A_pwl=eta_pwl ** (-n_pwl), n_pwl=n_pwl, Q_pwl=0.0, V_pwl=0.0, C=C,
                fric=math.radians(fric_deg), dil=0.0, Slim=1e50, eta_vp=0.0, n_vp=1.0,
                beta=1e-50)
    weak = dict(base, C=C * C_weak_ratio)
    return [base, weak]
