"""Shared test set-up: import paths, the `gpu` marker, common builders."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "mdoodz7.0_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and libmdgpu.so")


@pytest.fixture(scope="session")
def oracle_mod():
    import oracle
    oracle.lib()
    return oracle


def solvi_bcs(g, S):
    """Analytic Dirichlet tags/values of the SolVi test (SolViBenchmarkTests.cpp:104-160)."""
    Nx, Nz = g.Nx, g.Nz
    for l in range(Nz + 1):
        for k in range(Nx):
            c = k + l * Nx
            if k in (0, Nx - 1) and 0 < l < Nz:
                vx, _, _ = S.solvi_analytic(g.xg[k], g.zvx[l])
                g.BCu_t[c] = 0; g.BCu_v[c] = vx
            elif l == 0 or l == Nz:
                zz = g.zvx[l] + (g.dx / 2 if l == 0 else -g.dx / 2)
                vx, _, _ = S.solvi_analytic(g.xg[k], zz)
                g.BCu_t[c] = 11; g.BCu_v[c] = 2 * vx
    for l in range(Nz):
        for k in range(Nx + 1):
            c = k + l * (Nx + 1)
            if k in (0, Nx):
                xx = g.xvz[k] + (g.dx / 2 if k == 0 else -g.dx / 2)
                _, vz, _ = S.solvi_analytic(xx, g.zg[l])
                g.BCv_t[c] = 11; g.BCv_v[c] = 2 * vz
            elif l in (0, Nz - 1):
                _, vz, _ = S.solvi_analytic(g.xvz[k], g.zg[l])
                g.BCv_t[c] = 0; g.BCv_v[c] = vz
    return g
