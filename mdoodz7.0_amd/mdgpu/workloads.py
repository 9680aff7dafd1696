"""Synthetic workloads for the benchmark and the large-size tests.

The 2047^2 ellipse workload: unit square, 256 seeded ellipses (full axes
5-60 cells, angle U[0, pi), log10 viscosity U[-3, 3] against a matrix of 0,
density 3000 +- 300), nper x nper jittered markers per cell.  Phases are found
per ellipse only over the cells of its bounding box, using the marker order of
setups.put_markers (cell column outer, cell row inner, nper^2 markers per
cell), so building 37.7M markers takes seconds; the result is identical to
setups.ellipse_phase.
"""
import math

import numpy as np

from . import setups as S


def ellipse_phase_cells(x, z, ell, ncx, ncz, nper2, xmin, zmin, dx, dz):
    ph = np.zeros(x.size, np.int32)
    for e in range(len(ell["xc"])):
        c, s = math.cos(ell["theta"][e]), math.sin(ell["theta"][e])
        r = max(ell["a"][e], ell["b"][e])
        j0 = max(int((ell["xc"][e] - r - xmin) / dx) - 1, 0)
        j1 = min(int((ell["xc"][e] + r - xmin) / dx) + 1, ncx - 1)
        i0 = max(int((ell["zc"][e] - r - zmin) / dz) - 1, 0)
        i1 = min(int((ell["zc"][e] + r - zmin) / dz) + 1, ncz - 1)
        if j1 < j0 or i1 < i0:
            continue
        for j in range(j0, j1 + 1):
            sl = slice((j * ncz + i0) * nper2, (j * ncz + i1 + 1) * nper2)
            xr = x[sl] - ell["xc"][e]
            zr = z[sl] - ell["zc"][e]
            u = (c * xr + s * zr) / ell["a"][e]
            v = (-s * xr + c * zr) / ell["b"][e]
            p = ph[sl]
            p[(u * u + v * v) <= 1.0] = e + 1
    return ph


def ellipse_workload(ncells, nper, seed, n_ell=256):
    Nx = ncells + 1
    gc = S.grid_coords(Nx, Nx, 0.0, 1.0, 0.0, 1.0)
    x, z = S.put_markers(gc["xg"], gc["zg"], gc["dx"], gc["dz"], nper, nper, seed=seed)
    ell = S.random_ellipses(seed, n_ell, gc["dx"], 0.0, 1.0, 0.0, 1.0)
    ph = ellipse_phase_cells(x, z, ell, ncells, ncells, nper * nper, 0.0, 0.0, gc["dx"], gc["dz"])
    eta_table = 10.0 ** np.concatenate([[0.0], ell["logeta"]])
    rho_table = np.concatenate([[3000.0], ell["rho"]])
    return dict(Nx=Nx, grid=gc, x=x, z=z, phase=ph, ell=ell, eta_table=eta_table,
                rho_table=rho_table)
