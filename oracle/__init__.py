"""CPU FP64 oracle for the MDOODZ hot path (ctypes wrapper over liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.  Each wrapped C
function cites the reference file:line it restates (see mdoracle.h).
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_bp_nd = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_dp = C.POINTER(C.c_double)
_bp = C.POINTER(C.c_int8)


class OrcGrid(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Nz", C.c_int), ("dx", C.c_double), ("dz", C.c_double),
                ("xg", _dp), ("zg", _dp), ("xc", _dp), ("zc", _dp), ("xvz", _dp), ("zvx", _dp),
                ("BCu_t", _bp), ("BCv_t", _bp), ("BCp_t", _bp), ("BCg_t", _bp),
                ("BCu_v", _dp), ("BCv_v", _dp)]


class KillerOpts(C.Structure):
    _fields_ = [("penalty", C.c_double), ("celvol", C.c_double), ("max_its_PH", C.c_int),
                ("rel_tol_KSP", C.c_double), ("max_its_KSP", C.c_int),
                ("lin_abs_div", C.c_double), ("lin_rel_div", C.c_double),
                ("lin_abs_mom", C.c_double), ("lin_rel_mom", C.c_double),
                ("its_PH", C.c_int), ("its_KSP_tot", C.c_int),
                ("maxdiv", C.c_double), ("maxdiv0", C.c_double),
                ("maxru", C.c_double), ("maxru0", C.c_double), ("status", C.c_int)]


LAW_TAIL = ["m_pwl", "r_pwl", "f_pwl", "a_pwl"] + \
    [k + "_lin" for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")] + \
    [k + "_gbs" for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")] + \
    ["G_exp", "S_exp", "Q_exp", "E_exp", "n_exp", "gs_ref"]


class OrcPhase(C.Structure):
    """One phase of mat_prop (mdoodz.h:158-170): the members ViscosityConcise reads."""
    _fields_ = [("cstv", C.c_int), ("pwlv", C.c_int), ("plast", C.c_int),
                ("eta0", C.c_double), ("pref_pwl", C.c_double), ("F_pwl", C.c_double),
                ("A_pwl", C.c_double), ("n_pwl", C.c_double), ("Q_pwl", C.c_double),
                ("V_pwl", C.c_double), ("C", C.c_double), ("fric", C.c_double),
                ("dil", C.c_double), ("Slim", C.c_double), ("eta_vp", C.c_double),
                ("n_vp", C.c_double), ("beta_for_K", C.c_double),
                ("linv", C.c_int), ("gbsv", C.c_int), ("expv", C.c_int), ("reserved", C.c_int)] + \
        [(k, C.c_double) for k in LAW_TAIL]


class OrcRheoModel(C.Structure):
    _fields_ = [("step", C.c_int), ("elastic", C.c_int), ("dt", C.c_double), ("R", C.c_double),
                ("zeroC", C.c_double), ("min_eta", C.c_double), ("max_eta", C.c_double),
                ("eta_tol", C.c_double), ("Tiimin", C.c_double), ("scale_E", C.c_double),
                ("scale_eta", C.c_double), ("T_peierls", C.c_double), ("d_min", C.c_double)]


class OrcPhaseSoft(C.Structure):
    _fields_ = [(k, C.c_double) for k in ("G", "bet", "C", "C_end", "phi", "phi_end", "psi",
                                           "psi_end", "pls_start", "pls_end", "H_c")] + \
        [(k, C.c_int) for k in ("phi_soft", "psi_soft", "coh_soft")]


def phase(**kw):
    """OrcPhase with reference-like defaults (constant viscosity 1, no plasticity)."""
    d = dict(cstv=1, pwlv=0, plast=0, eta0=1.0, pref_pwl=1.0, F_pwl=1.0, A_pwl=1.0, n_pwl=1.0,
             Q_pwl=0.0, V_pwl=0.0, C=1e50, fric=0.0, dil=0.0, Slim=1e50, eta_vp=0.0, n_vp=1.0,
             beta_for_K=1e-50, gs_ref=2e-3)
    d.update(kw)
    return OrcPhase(**d)


def rheo_model(**kw):
    d = dict(step=1, elastic=0, dt=1.0, R=8.314, zeroC=0.0, min_eta=1e-50, max_eta=1e50,
             eta_tol=1e-13, Tiimin=1e-50, scale_E=1.0, scale_eta=1.0, T_peierls=1473.15, d_min=1e-13)
    d.update(kw)
    return OrcRheoModel(**d)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        gp = C.POINTER(OrcGrid)
        L.orc_grid_coords.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                      C.c_double, _dp, _dp, f64p, f64p, f64p, f64p, f64p, f64p]
        L.orc_bc_pure_shear.argtypes = [gp, C.c_double]
        L.orc_eqn_numbering.argtypes = [gp, i32p, i32p, i32p, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
        L.orc_eqn_numbering.restype = C.c_int
        L.orc_marker_index.argtypes = [gp, C.c_int, C.c_int, f64p, f64p, i32p, i32p]
        L.orc_p2mastah.argtypes = [gp, C.c_int, f64p, f64p, i32p, f64p, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, f64p,
                                   C.c_int, C.c_void_p, C.c_int, C.c_int]
        L.orc_c2v.argtypes = [gp, C.c_int, f64p, f64p]
        L.orc_v2p.argtypes = [gp, f64p, f64p, C.c_int, C.c_int, f64p, f64p, i32p, f64p, f64p]
        L.orc_g2p_centroids.argtypes = [gp, f64p, C.c_int, f64p, f64p, i32p, f64p]
        L.orc_g2p_vertices.argtypes = [gp, f64p, C.c_int, f64p, f64p, i32p, f64p]
        L.orc_update_particle_stress.argtypes = [gp] + [f64p] * 6 + [C.c_int, f64p, f64p, i32p, f64p,
                                                                     f64p, f64p]
        L.orc_update_particle_pressure.argtypes = [gp, f64p, f64p, f64p, C.c_int, f64p, f64p, i32p,
                                                   f64p]
        L.orc_advect_rk4.argtypes = [gp, f64p, f64p, C.c_double, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int, f64p, f64p, i32p, C.c_void_p,
                                     C.c_void_p]
        L.orc_apply_bc.argtypes = [gp, f64p, f64p]
        L.orc_exz_n.argtypes = [gp, f64p, f64p]
        L.orc_strain_rates.argtypes = [gp, C.c_int, f64p, f64p, f64p, f64p, f64p, f64p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_eval_rhs.argtypes = [gp, C.c_double, C.c_double, f64p, f64p, f64p, f64p]
        L.orc_residual.argtypes = [gp, i32p, i32p, i32p, C.c_int, f64p, f64p, f64p, f64p,
                                   f64p, f64p, f64p, C.c_void_p, C.c_void_p, f64p, f64p]
        L.orc_assemble_picard.argtypes = [gp, i32p, i32p, i32p, C.c_int, C.c_int, f64p, f64p,
                                          f64p, i32p, i32p, f64p, i32p, i32p, f64p, i32p,
                                          i32p, f64p]
        L.orc_assemble_jacobian.argtypes = [gp, i32p, i32p, i32p, C.c_int, C.c_int, f64p, f64p,
                                            C.c_double, i32p, i32p, f64p, i32p, i32p, f64p]
        L.orc_assemble_jacobian.restype = None
        L.orc_killer_solve.argtypes = [C.c_int, C.c_int, i32p, i32p, f64p, i32p, i32p, f64p,
                                       i32p, i32p, f64p, f64p, f64p, f64p, f64p,
                                       C.POINTER(KillerOpts)]
        L.orc_killer_solve.restype = C.c_int
        L.orc_killer_solve_newton.argtypes = [C.c_int, C.c_int] + [i32p, i32p, f64p] * 5 + \
            [f64p, f64p, f64p, f64p, C.POINTER(KillerOpts)]
        L.orc_killer_solve_newton.restype = C.c_int
        L.orc_update_particle_stress_subgrid.argtypes = [gp, C.c_double, f64p, C.c_int] + [f64p] * 7 + \
            [C.c_int, f64p, f64p, i32p, f64p, f64p, f64p]
        L.orc_update_particle_stress_subgrid.restype = None
        L.orc_update_particle_pressure_subgrid.argtypes = [gp, C.c_double, f64p, f64p, f64p, f64p, f64p, C.c_int,
                                                           f64p, f64p, i32p, f64p]
        L.orc_update_particle_pressure_subgrid.restype = None
        L.orc_reseed.argtypes = [gp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, f64p, f64p, i32p, i32p, i32p,
                                 C.POINTER(C.c_int)]
        L.orc_reseed.restype = C.c_int
        L.orc_energy_matrix.argtypes = [gp, C.c_double, f64p, f64p, f64p, f64p, _bp_nd, f64p, i32p, i32p, f64p,
                                        f64p]
        L.orc_energy_matrix.restype = None
        L.orc_energy_step.argtypes = [gp, C.c_double, f64p, f64p, C.c_void_p, C.c_void_p, C.c_void_p, f64p, f64p,
                                      _bp_nd, f64p, C.c_int, C.c_double, f64p]
        L.orc_energy_step.restype = C.c_int
        L.orc_accumulated_strain.argtypes = [gp, C.c_double, C.c_void_p, C.c_void_p, C.c_int, f64p, f64p, i32p,
                                             C.c_void_p]
        L.orc_accumulated_strain.restype = C.c_int
        L.orc_rotate_stress.argtypes = [C.c_int, C.c_double, C.c_int, i32p, f64p, f64p, f64p, f64p]
        L.orc_rotate_stress.restype = None
        L.orc_set_band_order.argtypes = [gp, i32p, i32p, C.c_int]
        L.orc_viscosity_concise.argtypes = [C.POINTER(OrcPhase), C.POINTER(OrcRheoModel)] + \
            [C.c_double] * 12 + [_dp, _dp, _dp, C.POINTER(C.c_int)]
        L.orc_viscosity_concise.restype = C.c_double
        L.orc_rheology_nodes.argtypes = [C.POINTER(OrcRheoModel), C.POINTER(OrcPhase), C.c_int,
                                         C.c_int] + [C.c_void_p] * 16 + [C.c_void_p, C.c_int,
                                         C.c_int] + [C.c_void_p] * 7
        L.orc_rheology_nodes.restype = C.c_int
        L.orc_tangent_nodes.argtypes = L.orc_rheology_nodes.argtypes[:23] + [C.c_void_p]
        L.orc_tangent_nodes.restype = C.c_int
        L.orc_material_grid.argtypes = [C.POINTER(OrcPhaseSoft), C.c_int, C.c_int] + \
            [C.c_void_p] * 3 + [C.c_int, C.c_int] + [C.c_void_p] * 5
        _LIB = L
    return _LIB


def _ptr(a, t=_dp):
    return a.ctypes.data_as(t)


class Grid:
    """Staggered grid + boundary tags in the reference layout (GridRoutines.c:246-330)."""

    def __init__(self, Nx, Nz, xmin, xmax, zmin, zmax):
        self.Nx, self.Nz = Nx, Nz
        self.xmin, self.xmax, self.zmin, self.zmax = xmin, xmax, zmin, zmax
        self.xg = np.zeros(Nx); self.zg = np.zeros(Nz)
        self.xc = np.zeros(Nx - 1); self.zc = np.zeros(Nz - 1)
        self.xvz = np.zeros(Nx + 1); self.zvx = np.zeros(Nz + 1)
        dx, dz = C.c_double(), C.c_double()
        lib().orc_grid_coords(Nx, Nz, xmin, zmin, xmax, zmax, C.byref(dx), C.byref(dz),
                              self.xg, self.zg, self.xc, self.zc, self.xvz, self.zvx)
        self.dx, self.dz = dx.value, dz.value
        self.BCu_t = np.full(Nx * (Nz + 1), -1, np.int8)
        self.BCv_t = np.full((Nx + 1) * Nz, -1, np.int8)
        self.BCp_t = np.full((Nx - 1) * (Nz - 1), -1, np.int8)
        self.BCg_t = np.full(Nx * Nz, -1, np.int8)
        self.BCu_v = np.zeros(Nx * (Nz + 1))
        self.BCv_v = np.zeros((Nx + 1) * Nz)
        self._c = None

    @property
    def c(self):
        g = OrcGrid(self.Nx, self.Nz, self.dx, self.dz, _ptr(self.xg), _ptr(self.zg),
                    _ptr(self.xc), _ptr(self.zc), _ptr(self.xvz), _ptr(self.zvx),
                    _ptr(self.BCu_t, _bp), _ptr(self.BCv_t, _bp), _ptr(self.BCp_t, _bp),
                    _ptr(self.BCg_t, _bp), _ptr(self.BCu_v), _ptr(self.BCv_v))
        self._c = g
        return C.byref(g)

    def pure_shear(self, ebg):
        lib().orc_bc_pure_shear(self.c, ebg)
        return self

    def numbering(self):
        eu = np.zeros(self.Nx * (self.Nz + 1), np.int32)
        ev = np.zeros((self.Nx + 1) * self.Nz, np.int32)
        ep = np.zeros((self.Nx - 1) * (self.Nz - 1), np.int32)
        nm, nc = C.c_int(), C.c_int()
        neq = lib().orc_eqn_numbering(self.c, eu, ev, ep, C.byref(nm), C.byref(nc))
        return eu, ev, ep, nm.value, nc.value, neq

    def node_count(self, centroid):
        return {1: (self.Nx - 1) * (self.Nz - 1), 0: self.Nx * self.Nz,
                -1: self.Nx * (self.Nz + 1), -2: (self.Nx + 1) * self.Nz}[centroid]


def marker_index(g, centroid, x, z):
    ip = np.zeros(len(x), np.int32); jp = np.zeros(len(x), np.int32)
    lib().orc_marker_index(g.c, centroid, len(x), np.ascontiguousarray(x),
                           np.ascontiguousarray(z), ip, jp)
    return ip, jp


def v2p(g, u, v, x, z, phase, conserv=0):
    """ParticleRoutines.c:63-256: (vx, vz) at the markers."""
    n = len(x)
    vx, vz = np.zeros(n), np.zeros(n)
    lib().orc_v2p(g.c, np.ascontiguousarray(u, np.float64), np.ascontiguousarray(v, np.float64),
                  conserv, n, np.ascontiguousarray(x, np.float64), np.ascontiguousarray(z, np.float64),
                  np.ascontiguousarray(phase, np.int32), vx, vz)
    return vx, vz


def g2p_centroids(g, F, x, z, phase):
    """ParticleRoutines.c:1332-1580: a centre field at the markers."""
    out = np.zeros(len(x))
    lib().orc_g2p_centroids(g.c, np.ascontiguousarray(F, np.float64), len(x),
                            np.ascontiguousarray(x, np.float64), np.ascontiguousarray(z, np.float64),
                            np.ascontiguousarray(phase, np.int32), out)
    return out


def g2p_vertices(g, F, x, z, phase):
    """ParticleRoutines.c:1223-1325: a vertex field at the markers."""
    out = np.zeros(len(x))
    lib().orc_g2p_vertices(g.c, np.ascontiguousarray(F, np.float64), len(x),
                           np.ascontiguousarray(x, np.float64), np.ascontiguousarray(z, np.float64),
                           np.ascontiguousarray(phase, np.int32), out)
    return out


def update_particle_stress(g, sxxd, szzd, sxz, sxxd0, szzd0, sxz0, x, z, phase, msxxd, mszzd, msxz):
    """RheologyParticles.c:1725-2082 (no subgrid stress diffusion, no
    rotation): returns the new marker stresses."""
    f = lambda a: np.ascontiguousarray(a, np.float64)
    out = [np.array(a, np.float64) for a in (msxxd, mszzd, msxz)]
    lib().orc_update_particle_stress(g.c, f(sxxd), f(szzd), f(sxz), f(sxxd0), f(szzd0), f(sxz0), len(x),
                                     f(x), f(z), np.ascontiguousarray(phase, np.int32), *out)
    return out


def update_particle_pressure(g, p_in, p0_n, x, z, phase, P):
    """RheologyParticles.c:1621-1704 (no subgrid diffusion): returns (new P, dp)."""
    P = np.array(P, np.float64)
    dp = np.zeros(len(p_in))
    lib().orc_update_particle_pressure(g.c, np.ascontiguousarray(p_in, np.float64),
                                       np.ascontiguousarray(p0_n, np.float64), dp, len(x),
                                       np.ascontiguousarray(x, np.float64),
                                       np.ascontiguousarray(z, np.float64),
                                       np.ascontiguousarray(phase, np.int32), P)
    return P, dp


STRAIN_RATES = ("eII_pl", "eII_pwl", "eII_exp", "eII_lin", "eII_gbs", "eII_el", "eII_cst")
STRAINS = ("strain", "strain_el", "strain_pl", "strain_pl_vol", "strain_pwl", "strain_exp", "strain_lin",
           "strain_gbs")


def accumulated_strain(g, dt, x, z, phase, rates, strains, div_u_pl=None):
    """RheologyParticles.c:346-649: returns updated copies of the marker
    strains (dict), or raises on a negative increment."""
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
    keep = [f(rates.get(k)) for k in STRAIN_RATES]
    out = {k: np.array(v, np.float64) for k, v in strains.items()}
    e = (C.c_void_p * 7)(*[None if a is None else a.ctypes.data for a in keep])
    st = (C.c_void_p * 8)(*[out[k].ctypes.data if k in out else None for k in STRAINS])
    dp = f(div_u_pl)
    bad = lib().orc_accumulated_strain(g.c, dt, C.cast(e, C.c_void_p), None if dp is None else dp.ctypes.data, len(x),
                                       f(x), f(z), np.ascontiguousarray(phase, np.int32), C.cast(st, C.c_void_p))
    if bad >= 0:
        raise ValueError("negative strain increment at cell %d" % bad)
    return out


def update_particle_stress_subgrid(g, dt, G_phase, sxxd, szzd, sxz, sxxd0, szzd0, sxz0, eta_phys_s, x, z, phase,
                                   msxxd, mszzd, msxz, stencil=1):
    """UpdateParticleStress, subgrid_diffusion = 2 (RheologyParticles.c:1811-1895): returns new
    (msxxd, mszzd, msxz)."""
    f = lambda a: np.ascontiguousarray(a, np.float64)
    out = [np.array(a, np.float64) for a in (msxxd, mszzd, msxz)]
    lib().orc_update_particle_stress_subgrid(g.c, dt, f(G_phase), int(stencil), f(sxxd), f(szzd), f(sxz), f(sxxd0),
                                             f(szzd0), f(sxz0), f(eta_phys_s), len(x), f(x), f(z),
                                             np.ascontiguousarray(phase, np.int32), *out)
    return out


def update_particle_pressure_subgrid(g, dt, beta_phase, p_in, p0_n, eta_phys_n, x, z, phase, P):
    """UpdateParticlePressure, subgrid_diffusion >= 2 (RheologyParticles.c:1621-1695): returns
    (new P, dp)."""
    f = lambda a: np.ascontiguousarray(a, np.float64)
    out = np.array(P, np.float64)
    dp = np.zeros((g.Nx - 1) * (g.Nz - 1))
    lib().orc_update_particle_pressure_subgrid(g.c, dt, f(beta_phase), f(p_in), f(p0_n), f(eta_phys_n), dp, len(x),
                                               f(x), f(z), np.ascontiguousarray(phase, np.int32), out)
    return out, dp


def reseed(g, x, z, phase, min_part_cell=16, capacity=None, xmin=None, xmax=None):
    """CountPartCell_OLD re-seeding (ParticleRoutines.c:3224-4050).  Returns
    (x, z, phase) of length n_out, the slots and sources of the new markers,
    and the status (number of new markers, or -1 / -2 as orc_reseed)."""
    n = len(x)
    cap = capacity if capacity is not None else n + 4 * g.Nx * g.Nz + 64
    X = np.zeros(cap); Z = np.zeros(cap); P = np.full(cap, -1, np.int32)
    X[:n] = x; Z[:n] = z; P[:n] = phase
    nf = 4 * (g.Nx + 1) * (g.Nz + 1)
    slot = np.zeros(nf, np.int32); src = np.zeros(nf, np.int32)
    n_out = C.c_int(n)
    rc = lib().orc_reseed(g.c, g.xg[0] if xmin is None else xmin, g.xg[-1] if xmax is None else xmax,
                          int(min_part_cell), n, cap, X, Z, P, slot, src, C.byref(n_out))
    m = max(rc, 0)
    return X[:n_out.value], Z[:n_out.value], P[:n_out.value], slot[:m], src[:m], rc


def energy_matrix(g, dt, rho_n, Cp, kx, kz, BCT_t, BCT_v):
    """ThermalRoutines.c:185-390: (scipy CSR matrix, boundary rhs)."""
    n = (g.Nx - 1) * (g.Nz - 1)
    Ap = np.zeros(n + 1, np.int32); Aj = np.zeros(5 * n, np.int32); Ax = np.zeros(5 * n); bbc = np.zeros(n)
    f = lambda a: np.ascontiguousarray(a, np.float64)
    lib().orc_energy_matrix(g.c, dt, f(rho_n), f(Cp), f(kx), f(kz), np.ascontiguousarray(BCT_t, np.int8), f(BCT_v),
                            Ap, Aj, Ax, bbc)
    import scipy.sparse as sp
    return sp.csr_matrix((Ax[:Ap[-1]], Aj[:Ap[-1]], Ap), shape=(n, n)), bbc


def energy_step(g, dt, rho_n, Cp, T, kx, kz, BCT_t, BCT_v, Qr=None, Wdiss=None, Ha=None, max_its=1000,
                rel_tol=1e-8):
    """EnergyDirectSolve + SolveThermalPCG: (new T, iterations)."""
    f = lambda a: None if a is None else np.ascontiguousarray(a, np.float64)
    Tn = np.array(T, np.float64)
    extra = [f(a) for a in (Qr, Wdiss, Ha)]
    its = lib().orc_energy_step(g.c, dt, f(rho_n), f(Cp), *[None if a is None else a.ctypes.data for a in extra],
                                f(kx), f(kz), np.ascontiguousarray(BCT_t, np.int8), f(BCT_v), int(max_its),
                                rel_tol, Tn)
    return Tn, its


def rotate_stress(mode, dt, phase, wxzm, sxxd, szzd, sxz):
    """RheologyParticles.c:1965-1988: rotated copies of the marker stresses."""
    out = [np.array(a, np.float64) for a in (sxxd, szzd, sxz)]
    lib().orc_rotate_stress(int(mode), dt, len(phase), np.ascontiguousarray(phase, np.int32),
                            np.ascontiguousarray(wxzm, np.float64), *out)
    return out


def advect_rk4(g, u, v, dt, x, z, phase, conserv=0, vx=None, vz=None):
    """AdvectionRoutines.c:347-410: returns new (x, z, phase, vx, vz); vx, vz
    keep their input values (zeros if not given) for markers that end
    inactive."""
    n = len(x)
    x, z = np.array(x, np.float64), np.array(z, np.float64)
    phase = np.array(phase, np.int32)
    vx = np.zeros(n) if vx is None else np.array(vx, np.float64)
    vz = np.zeros(n) if vz is None else np.array(vz, np.float64)
    lib().orc_advect_rk4(g.c, np.ascontiguousarray(u, np.float64), np.ascontiguousarray(v, np.float64),
                         dt, conserv, g.xmin, g.xmax, g.zmin, g.zmax, n, x, z, phase,
                         vx.ctypes.data, vz.ctypes.data)
    return x, z, phase, vx, vz


def p2mastah(g, x, z, phase, mat_prop, flag=0, avg=0, prop=0, centroid=1, stencil=1,
             periodic_x=0, node_type=None, node_field=None, nb_phases=0, mode=0, nthreads=1):
    """ParticleRoutines.c:2170-2665.  Returns node_field (prop=0) or phase_perc (prop=1)."""
    nn = g.node_count(centroid)
    if node_field is None:
        node_field = np.zeros(nn)
    perc = np.zeros(max(nb_phases, 1) * nn) if prop == 1 else None
    nt = node_type.ctypes.data if node_type is not None else None
    lib().orc_p2mastah(g.c, len(x), np.ascontiguousarray(x, np.float64),
                       np.ascontiguousarray(z, np.float64), np.ascontiguousarray(phase, np.int32),
                       np.ascontiguousarray(mat_prop, np.float64), flag, avg, prop, centroid,
                       stencil, periodic_x, nt, node_field, nb_phases,
                       perc.ctypes.data if perc is not None else None, mode, nthreads)
    if prop == 1:
        return perc.reshape(nb_phases, nn)
    return node_field


def c2v(g, cent, periodic_x=0):
    v = np.zeros(g.Nx * g.Nz)
    lib().orc_c2v(g.c, periodic_x, np.ascontiguousarray(cent), v)
    return v


def apply_bc(g, u, v):
    lib().orc_apply_bc(g.c, u, v)


def strain_rates(g, u, v, compressible=0):
    nc = (g.Nx - 1) * (g.Nz - 1); nv = g.Nx * g.Nz
    out = {k: np.zeros(nc) for k in ("div", "exxd", "ezzd")}
    out["exz"] = np.zeros(nv); out["wxz"] = np.zeros(nv)
    lib().orc_strain_rates(g.c, compressible, u, v, out["div"], out["exxd"], out["ezzd"],
                           out["exz"], out["wxz"].ctypes.data, None, None, None)
    return out


def exz_n(g, exz):
    out = np.zeros((g.Nx - 1) * (g.Nz - 1))
    lib().orc_exz_n(g.c, np.ascontiguousarray(exz), out)
    return out


def eval_rhs(g, gx, gz, rho_n, rho_s):
    rx = np.zeros(g.Nx * (g.Nz + 1)); rz = np.zeros((g.Nx + 1) * g.Nz)
    lib().orc_eval_rhs(g.c, gx, gz, rho_n, rho_s, rx, rz)
    return rx, rz


def residual(g, num, sxxd, szzd, sxz, p_corr, div, roger_x, roger_z, d=None):
    eu, ev, ep, nm, nc, neq = num
    F = np.zeros(neq)
    norms = np.zeros(3)
    lib().orc_residual(g.c, eu, ev, ep, nm, sxxd, szzd, sxz, p_corr, div, roger_x, roger_z,
                       None, d.ctypes.data if d is not None else None, F, norms)
    return F, norms


def assemble_picard(g, num, eta_n, eta_s):
    """CSR blocks A (nm x nm), B (nm x nc), C (nc x nm) of the Picard operator."""
    eu, ev, ep, nm, nc, neq = num
    D11 = np.ascontiguousarray(2.0 * eta_n); D33 = np.ascontiguousarray(eta_s)
    Ap = np.zeros(nm + 1, np.int32); Aj = np.zeros(9 * nm + 1, np.int32); Ax = np.zeros(9 * nm + 1)
    Bp = np.zeros(nm + 1, np.int32); Bj = np.zeros(2 * nm + 1, np.int32); Bx = np.zeros(2 * nm + 1)
    Cp = np.zeros(nc + 1, np.int32); Cj = np.zeros(4 * nc + 1, np.int32); Cx = np.zeros(4 * nc + 1)
    lib().orc_assemble_picard(g.c, eu, ev, ep, nm, nc, D11, D11, D33, Ap, Aj, Ax, Bp, Bj, Bx,
                              Cp, Cj, Cx)
    import scipy.sparse as sp
    A = sp.csr_matrix((Ax[:Ap[-1]], Aj[:Ap[-1]], Ap), shape=(nm, nm))
    B = sp.csr_matrix((Bx[:Bp[-1]], Bj[:Bp[-1]], Bp), shape=(nm, nc))
    Cm = sp.csr_matrix((Cx[:Cp[-1]], Cj[:Cp[-1]], Cp), shape=(nc, nm))
    return A, B, Cm


def assemble_jacobian(g, num, Dn, Ds, co=1.0):
    """StokesAssemblyDecoupled.c:219-1085,1786-2304: the Newton momentum
    blocks (AJ, BJ) as scipy CSR; Dn (8, cells), Ds (4, vertices); co =
    comp*out_of_plane (BuildJacobianOperatorDecoupled uses comp = 1)."""
    eu, ev, ep, nm, nc, neq = num
    Dn = np.ascontiguousarray(Dn, np.float64).reshape(-1)
    Ds = np.ascontiguousarray(Ds, np.float64).reshape(-1)
    Ap = np.zeros(nm + 1, np.int32); Aj = np.zeros(21 * nm + 1, np.int32); Ax = np.zeros(21 * nm + 1)
    Bp = np.zeros(nm + 1, np.int32); Bj = np.zeros(6 * nm + 1, np.int32); Bx = np.zeros(6 * nm + 1)
    lib().orc_assemble_jacobian(g.c, eu, ev, ep, nm, nc, Dn, Ds, float(co), Ap, Aj, Ax, Bp, Bj, Bx)
    import scipy.sparse as sp
    AJ = sp.csr_matrix((Ax[:Ap[-1]], Aj[:Ap[-1]], Ap), shape=(nm, nm))
    BJ = sp.csr_matrix((Bx[:Bp[-1]], Bj[:Bp[-1]], Bp), shape=(nm, nc))
    AJ.sum_duplicates()
    BJ.sum_duplicates()
    return AJ, BJ


def killer_solve(g, num, A, B, Cm, bu, bp, penalty=1e5, max_its_PH=10, rel_tol_KSP=1e-4,
                 max_its_KSP=100, lin_abs_div=1e-9, lin_rel_div=1e-5, lin_abs_mom=1e-9,
                 lin_rel_mom=1e-5, AJ=None, BJ=None):
    """Solvers.c:1054-1454 (PH + GCR(6), exact band-Cholesky preconditioner
    of the Picard Jts).  With the Newton blocks AJ, BJ (assemble_jacobian),
    GCR iterates on the Jacobian system as the reference does."""
    eu, ev, ep, nm, nc, neq = num
    lib().orc_set_band_order(g.c, eu, ev, nm)
    o = KillerOpts(penalty, g.dx * g.dz, max_its_PH, rel_tol_KSP, max_its_KSP,
                   lin_abs_div, lin_rel_div, lin_abs_mom, lin_rel_mom, 0, 0, 0, 0, 0, 0, 0)
    du = np.zeros(nm); dp = np.zeros(nc)

    def csr(M):
        return (np.ascontiguousarray(M.indptr, np.int32), np.ascontiguousarray(M.indices, np.int32),
                np.ascontiguousarray(M.data, np.float64))
    a, b, c = csr(A), csr(B), csr(Cm)
    if AJ is None:
        st = lib().orc_killer_solve(nm, nc, *a, *b, *c, np.ascontiguousarray(bu),
                                    np.ascontiguousarray(bp), du, dp, C.byref(o))
    else:
        ja, jb = csr(AJ.tocsr()), csr(BJ.tocsr())
        st = lib().orc_killer_solve_newton(nm, nc, *a, *b, *ja, *jb, *c, np.ascontiguousarray(bu),
                                           np.ascontiguousarray(bp), du, dp, C.byref(o))
    return du, dp, o, st


def viscosity_concise(ph, md, G, T, P, Exx, Ezz, Exz, fric=None, Coh=None, dil=None,
                      pref_mult=1.0, d=0.0, phi=0.0):
    """RheologyDensity.c:385-1080 on one (node, phase) at grain size d and
    melt fraction phi: returns (eta_dissipative, etaVE, Pcorr, Eii_pl,
    status)."""
    e = C.c_double(); pc = C.c_double(); epl = C.c_double(); st = C.c_int(0)
    eta = lib().orc_viscosity_concise(C.byref(ph), C.byref(md), G, T, P, d, phi, Exx, Ezz, Exz,
                                      ph.fric if fric is None else fric,
                                      ph.C if Coh is None else Coh,
                                      ph.dil if dil is None else dil, pref_mult,
                                      C.byref(e), C.byref(pc), C.byref(epl), C.byref(st))
    return eta, e.value, pc.value, epl.value, st.value


def rheology_nodes(md, phases, perc, G, T, P, exx, ezz, exz, sxx0=None, szz0=None, sxz0=None,
                   fric=None, Coh=None, dil=None, pref=None, tag=None, centroid=1, average=0, d=None,
                   phi=None):
    """NonNewtonianViscosityGrid (RheologyDensity.c:1131-1426) on one node set.
    perc has shape (nph, n).  Returns dict of eta, eta_phys, sxx, szz, sxz, p_corr,
    eII_pl and the status (-1 when a local iteration failed)."""
    perc = np.ascontiguousarray(perc, np.float64)
    nph, n = perc.shape
    arr = (C.c_double * 1)
    keep = []

    def p(a, dt=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), (n,)))
        keep.append(a)
        return a.ctypes.data
    z = np.zeros(n)
    outs = {k: np.zeros(n) for k in ("eta", "eta_phys", "sxx", "szz", "sxz", "p_corr", "eII_pl")}
    ph = (OrcPhase * nph)(*phases)
    st = lib().orc_rheology_nodes(C.byref(md), ph, nph, n, perc.ctypes.data, p(G), p(T), p(P),
                                  p(exx), p(ezz), p(exz), p(z if sxx0 is None else sxx0),
                                  p(z if szz0 is None else szz0), p(z if sxz0 is None else sxz0),
                                  p(fric), p(Coh), p(dil), p(pref), p(d), p(phi), p(tag, np.int8),
                                  centroid, average, *[outs[k].ctypes.data for k in
                                             ("eta", "eta_phys", "sxx", "szz", "sxz", "p_corr",
                                              "eII_pl")])
    outs["status"] = st
    return outs


def material_grid(phases, perc, strain_pl=None, tag=None, centroid=1, style=0):
    """ShearModCompExpGrid + CohesionFrictionDilationGrid (RheologyDensity.c:
    1432-1730) on one node set; phases are OrcPhaseSoft, perc (nph, n)."""
    perc = np.ascontiguousarray(perc, np.float64)
    nph, n = perc.shape
    out = {k: np.zeros(n) for k in ("mu", "bet", "C", "fric", "dil")}
    sp = None if strain_pl is None else np.ascontiguousarray(strain_pl, np.float64)
    tg = None if tag is None else np.ascontiguousarray(tag, np.int8)
    arr = (OrcPhaseSoft * nph)(*phases)
    lib().orc_material_grid(arr, nph, n, perc.ctypes.data,
                            None if sp is None else sp.ctypes.data,
                            None if tg is None else tg.ctypes.data, centroid, style,
                            *[out[k].ctypes.data for k in ("mu", "bet", "C", "fric", "dil")])
    return out


def tangent_nodes(md, phases, perc, G, T, P, exx, ezz, exz, sxx0=None, szz0=None, sxz0=None,
                  fric=None, Coh=None, dil=None, pref=None, tag=None, centroid=1, average=0, d=None,
                  phi=None):
    """Newton tangent coefficients (StokesRoutines.c:250-375 with
    FD_Jacobian.c:48-323): array (8, n) on centres (D11..D14, D21..D24) or
    (4, n) on vertices (D31..D34)."""
    perc = np.ascontiguousarray(perc, np.float64)
    nph, n = perc.shape
    keep = []

    def p(a, dt=np.float64):
        if a is None:
            return None
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(a, dt), (n,)))
        keep.append(a)
        return a.ctypes.data
    z = np.zeros(n)
    D = np.zeros(((8 if centroid else 4), n))
    ph = (OrcPhase * nph)(*phases)
    st = lib().orc_tangent_nodes(C.byref(md), ph, nph, n, perc.ctypes.data, p(G), p(T), p(P),
                                 p(exx), p(ezz), p(exz), p(z if sxx0 is None else sxx0),
                                 p(z if szz0 is None else szz0), p(z if sxz0 is None else sxz0),
                                 p(fric), p(Coh), p(dil), p(pref), p(d), p(phi), p(tag, np.int8),
                                 centroid, average, D.ctypes.data)
    if st:
        raise RuntimeError("local iterations failed in the tangent")
    return D
