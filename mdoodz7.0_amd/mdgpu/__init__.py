"""Python host side of the MI355X hot path (ctypes over libmdgpu.so).

The names mirror the reference call sites: Context.p2mastah is P2Mastah
(ParticleRoutines.c:2170), Context.p2mastah_fused is P2Mastah_Fused
(ParticleRoutines.c:1895), Context.killer_solver replaces KillerSolver
(Solvers.c:1054) at the lin_solver switch (StokesRoutines.c:845-848).
Arguments keep their reference meaning; failures raise MdgError carrying the
library's message instead of calling exit().

Arrays may be numpy arrays (host) or DeviceArray objects (device memory
allocated through the library's own HIP runtime); device arrays are used in
place.  There is no CPU fallback: without the built library or a GPU every
call raises.
"""
import ctypes as C
import os
import re

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "libmdgpu.so")
HEADER_PATH = os.path.join(os.path.dirname(os.path.dirname(_HERE)), "include", "mdgpu.h")
_LIB = None

MDG_SRC_SORTED = 0x100


class MdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mdgpu error %d: %s" % (code, msg))
        self.code = code


class GridDesc(C.Structure):
    _fields_ = [("Nx", C.c_int), ("Nz", C.c_int), ("xmin", C.c_double), ("xmax", C.c_double),
                ("zmin", C.c_double), ("zmax", C.c_double), ("periodic_x", C.c_int)]


class P2GField(C.Structure):
    _fields_ = [("src", C.c_void_p), ("n_src", C.c_int), ("dst", C.c_void_p),
                ("node_type", C.c_void_p), ("flag", C.c_int), ("avg", C.c_int)]


def declared_functions():
    """Names of every function declared in include/mdgpu.h."""
    txt = open(HEADER_PATH).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(mdg_[a-z0-9_]+)\s*\(", txt)))


def lib():
    """Load libmdgpu.so (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libmdgpu.so not built: run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.mdg_create.argtypes = [C.POINTER(GridDesc), C.POINTER(vp)]
        L.mdg_destroy.argtypes = [vp]
        L.mdg_destroy.restype = None
        L.mdg_last_error.argtypes = [vp]
        L.mdg_last_error.restype = C.c_char_p
        L.mdg_grid_coords.argtypes = [vp] + [vp] * 6
        L.mdg_set_bc.argtypes = [vp] * 7
        L.mdg_markers_set.argtypes = [vp, C.c_int, vp, vp, vp]
        L.mdg_marker_index.argtypes = [vp, C.c_int, vp, vp]
        L.mdg_markers_gather.argtypes = [vp, vp, vp]
        L.mdg_p2mastah.argtypes = [vp, vp, C.c_int, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, vp]
        L.mdg_p2mastah_fused.argtypes = [vp, C.c_int, C.c_int, C.POINTER(P2GField), C.c_int]
        L.mdg_last_kernel_ms.argtypes = [vp, C.POINTER(C.c_double)]
        L.mdg_last_weights_ms.argtypes = [vp, C.POINTER(C.c_double)]
        L.mdg_dev_alloc.argtypes = [vp, C.c_ulonglong, C.POINTER(vp)]
        L.mdg_dev_free.argtypes = [vp, vp]
        L.mdg_memcpy.argtypes = [vp, vp, vp, C.c_ulonglong]
        L.mdg_dev_scale.argtypes = [vp, C.c_longlong, C.c_double, vp, vp]
        L.mdg_v2p.argtypes = [vp, vp, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.mdg_advect_rk4.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        L.mdg_g2p_centroids.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.mdg_g2p_vertices.argtypes = [vp, vp, C.c_int, vp, vp, vp, vp]
        L.mdg_update_particle_stress.argtypes = [vp] * 7 + [C.c_int] + [vp] * 6
        L.mdg_update_particle_pressure.argtypes = [vp, vp, vp, vp, C.c_int, vp, vp, vp, vp]
        L.mdg_accumulated_strain.argtypes = [vp, C.c_double, vp, vp, C.c_int, vp, vp, vp, vp]
        L.mdg_rotate_particle_stress.argtypes = [vp, C.c_int, C.c_double, C.c_int, vp, vp, vp, vp, vp]
        L.mdg_update_particle_stress_subgrid.argtypes = [vp, C.c_double, vp, C.c_int, C.c_int] + [vp] * 7 + \
            [C.c_int] + [vp] * 6
        L.mdg_update_particle_pressure_subgrid.argtypes = [vp, C.c_double, vp, C.c_int] + [vp] * 4 + [C.c_int] + [vp] * 4
        L.mdg_count_part_cell.argtypes = [vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp, vp, C.POINTER(C.c_int),
                                          C.POINTER(C.c_int)]
        L.mdg_energy_step.argtypes = [vp, C.c_double, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_int, C.c_double, vp,
                                      C.POINTER(C.c_int)]
        _extra_signatures(L)
        _LIB = L
    return _LIB


class SolverOpts(C.Structure):
    """mdg_solver_opts (include/mdgpu.h)."""
    _fields_ = [("penalty", C.c_double), ("rel_tol", C.c_double), ("max_its", C.c_int),
                ("restart", C.c_int), ("inner_tol", C.c_double), ("inner_max_its", C.c_int),
                ("cheb_degree", C.c_int), ("coarse_cells", C.c_int), ("precision", C.c_int),
                ("precond", C.c_int), ("cycle", C.c_int), ("eta_avg", C.c_int),
                ("check_every", C.c_int), ("schur", C.c_int), ("poisson_tol", C.c_double),
                ("poisson_max_its", C.c_int), ("smoother", C.c_int),
                ("cheb_degree_coarse", C.c_int), ("cheb_range", C.c_double), ("coarsen4", C.c_int),
                ("graphs", C.c_int),
                ("its", C.c_int),
                ("inner_its", C.c_int),
                ("poisson_its", C.c_int),
                ("levels", C.c_int), ("status", C.c_int), ("rel_res", C.c_double),
                ("time_ms", C.c_double), ("guess_used", C.c_int), ("guess_gain", C.c_double),
                ("method", C.c_int), ("max_its_PH", C.c_int), ("rel_tol_KSP", C.c_double),
                ("max_its_KSP", C.c_int), ("lin_abs_mom", C.c_double), ("lin_rel_mom", C.c_double),
                ("lin_abs_div", C.c_double), ("lin_rel_div", C.c_double), ("pc_tol", C.c_double),
                ("max_mom", C.c_double), ("max_mom0", C.c_double), ("max_div", C.c_double),
                ("max_div0", C.c_double), ("its_PH", C.c_int), ("its_KSP", C.c_int),
                ("smoother_ms", C.c_double), ("smoother_launches", C.c_int), ("patch_levels", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


MDG_PREC_FP64, MDG_PREC_MIXED = 0, 1
MDG_ERR_DIVERGED = -6
MDG_ERR_INCOMPRESSIBLE = -8


LAW_TAIL = ("m_pwl", "r_pwl", "f_pwl", "a_pwl") + \
    tuple(k + "_lin" for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")) + \
    tuple(k + "_gbs" for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")) + \
    ("G_exp", "S_exp", "Q_exp", "E_exp", "n_exp", "gs_ref")


class PhaseProp(C.Structure):
    """mdg_phase_prop: one phase of mat_prop for the implemented creep laws
    (constant, power law, diffusion, grain-boundary sliding, Peierls)."""
    _fields_ = [("cstv", C.c_int), ("pwlv", C.c_int), ("plast", C.c_int),
                ("eta0", C.c_double), ("pref_pwl", C.c_double), ("F_pwl", C.c_double),
                ("A_pwl", C.c_double), ("n_pwl", C.c_double), ("Q_pwl", C.c_double),
                ("V_pwl", C.c_double), ("C", C.c_double), ("fric", C.c_double),
                ("dil", C.c_double), ("Slim", C.c_double), ("eta_vp", C.c_double),
                ("n_vp", C.c_double), ("beta", C.c_double),
                ("linv", C.c_int), ("gbsv", C.c_int), ("expv", C.c_int), ("reserved", C.c_int)] + \
        [(k, C.c_double) for k in LAW_TAIL]


class RheoParams(C.Structure):
    """mdg_rheo_params: the params/scale fields ViscosityConcise reads
    (T_peierls = 0 leaves Peierls creep inactive)."""
    _fields_ = [("step", C.c_int), ("elastic", C.c_int), ("average", C.c_int),
                ("dt", C.c_double), ("R", C.c_double), ("zeroC", C.c_double),
                ("min_eta", C.c_double), ("max_eta", C.c_double), ("eta_tol", C.c_double),
                ("Tiimin", C.c_double), ("scale_E", C.c_double), ("scale_eta", C.c_double),
                ("T_peierls", C.c_double), ("d_min", C.c_double)]


RHEO_IN = ("perc", "G", "T", "P", "exx", "ezz", "exz", "sxx0", "szz0", "sxz0", "fric", "C",
           "dil", "pref", "d", "phi")
RHEO_OUT = ("eta", "eta_phys", "sxx", "szz", "sxz", "p_corr", "eII_pl")


class PhaseSoft(C.Structure):
    """mdg_phase_soft: moduli and softening parameters of one phase."""
    _fields_ = [(k, C.c_double) for k in ("G", "beta", "C", "C_end", "fric", "fric_end", "dil",
                                           "dil_end", "pls_start", "pls_end", "H_c")] + \
        [(k, C.c_int) for k in ("fric_soft", "dil_soft", "coh_soft")]


class RheoNodes(C.Structure):
    """mdg_rheo_nodes: grain size d and melt fraction phi follow the outputs."""
    _fields_ = [(k, C.c_void_p) for k in RHEO_IN[:14] + RHEO_OUT + RHEO_IN[14:]]


def _extra_signatures(L):
    """Signatures of the Stokes and solver entry points."""
    vp = C.c_void_p
    L.mdg_eqn_count.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.mdg_stokes_apply.argtypes = [vp] * 6
    L.mdg_apply_bc.argtypes = [vp, vp, vp]
    L.mdg_strain_rates.argtypes = [vp] * 7
    L.mdg_extract_solution.argtypes = [vp, vp, C.c_double, vp, vp, vp]
    L.mdg_eval_rhs.argtypes = [vp, C.c_double, C.c_double, vp, vp, vp, vp]
    L.mdg_stokes_residual.argtypes = [vp] * 11
    L.mdg_stokes_residual_static.argtypes = [vp] * 11
    L.mdg_killer_solver.argtypes = [vp, C.POINTER(SolverOpts), vp, vp, vp, vp, vp, vp]
    L.mdg_time_smoother.argtypes = [vp, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.mdg_c2v.argtypes = [vp, C.c_int, vp, vp]
    L.mdg_stokes_diag.argtypes = [vp] * 5
    L.mdg_tangent_grid.argtypes = [vp, C.POINTER(PhaseProp), C.c_int, C.POINTER(RheoParams),
                                   C.c_int, C.POINTER(RheoNodes), vp]
    L.mdg_jacobian_apply.argtypes = [vp] * 5
    L.mdg_newton_solver.argtypes = [vp, C.POINTER(SolverOpts)] + [vp] * 7
    L.mdg_material_grid.argtypes = [vp, C.POINTER(PhaseSoft), C.c_int, C.c_int, C.c_int] + [vp] * 7
    L.mdg_pcnet_param_count.argtypes = [vp, C.POINTER(C.c_longlong)]
    L.mdg_pcnet_set_params.argtypes = [vp, vp, C.c_longlong]
    L.mdg_pcnet_get_params.argtypes = [vp, vp, vp]
    L.mdg_pcnet_forward.argtypes = [vp, C.c_int, C.c_int, vp, vp]
    L.mdg_pcnet_grad.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, C.c_int, C.POINTER(C.c_double)]
    L.mdg_pcnet_adam.argtypes = [vp, C.c_float]
    L.mdg_pcnet_capture.argtypes = [vp, C.c_int]
    L.mdg_pcnet_capture_outer.argtypes = [vp, C.c_int]
    L.mdg_pcnet_captured.argtypes = [vp] + [C.POINTER(C.c_int)] * 3
    L.mdg_pcnet_get_pair.argtypes = [vp, C.c_int, vp, vp, vp]
    L.mdg_pcnet_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, vp]
    L.mdg_pcnet_train_residual.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_ulonglong, vp]
    L.mdg_strain_rates_nodes.argtypes = [vp] * 9
    L.mdg_viscosity_grid.argtypes = [vp, C.POINTER(PhaseProp), C.c_int, C.POINTER(RheoParams),
                                     C.c_int, C.POINTER(RheoNodes)]


def ptr(a):
    """Raw address of a numpy array or DeviceArray (None passes NULL)."""
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.p
    if isinstance(a, np.ndarray):
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError("arrays must be C-contiguous")
        return a.ctypes.data
    raise TypeError("expected a numpy array or DeviceArray, got %r" % type(a))


def checked(a, dtype, n, name, optional=False):
    """ptr(a) after checking that a holds exactly n elements of dtype (a
    DeviceArray may be longer).  The library reads and writes these arrays
    through raw pointers, so a wrong dtype or a short array would be misread
    or overrun without any error of its own."""
    if a is None:
        if optional:
            return None
        raise ValueError("%s is required" % name)
    dt = np.dtype(dtype)
    if isinstance(a, DeviceArray):
        if a.dtype != dt:
            raise ValueError("%s: expected %s, got %s" % (name, dt, a.dtype))
        if a.n < n:
            raise ValueError("%s: %d elements, need %d" % (name, a.n, n))
        return a.p
    if isinstance(a, np.ndarray):
        if a.dtype != dt:
            raise ValueError("%s: expected %s, got %s" % (name, dt, a.dtype))
        if a.size != n:
            raise ValueError("%s: %d elements, expected %d" % (name, a.size, n))
        return ptr(a)
    raise TypeError("%s: expected a numpy array or DeviceArray, got %r" % (name, type(a)))


def _flat(a):
    """A DeviceArray as it is, anything else as a contiguous float64 array."""
    return a if isinstance(a, DeviceArray) else np.ascontiguousarray(a, np.float64)


class DeviceArray:
    """A flat device array allocated with mdg_dev_alloc on a context."""

    def __init__(self, ctx, n, dtype=np.float64, host=None):
        self.ctx, self.n, self.dtype = ctx, int(n), np.dtype(dtype)
        h = C.c_void_p()
        ctx._check(ctx._L.mdg_dev_alloc(ctx.h, self.n * self.dtype.itemsize, C.byref(h)))
        self.p = h.value
        if host is not None:
            self.upload(host)

    def __len__(self):
        return self.n

    def upload(self, host):
        host = np.ascontiguousarray(host, self.dtype)
        if host.size != self.n:
            raise ValueError("upload: %d elements into a device array of %d" % (host.size, self.n))
        self.ctx._check(self.ctx._L.mdg_memcpy(self.ctx.h, self.p, host.ctypes.data, host.nbytes))
        return self

    def numpy(self):
        out = np.empty(self.n, self.dtype)
        self.ctx._check(self.ctx._L.mdg_memcpy(self.ctx.h, out.ctypes.data, self.p, out.nbytes))
        return out

    def copy_from(self, other):
        """Device-to-device copy of another DeviceArray of the same size."""
        if other.n != self.n or other.dtype != self.dtype:
            raise ValueError("copy_from: %d x %s into %d x %s" % (other.n, other.dtype, self.n, self.dtype))
        self.ctx._check(self.ctx._L.mdg_memcpy(self.ctx.h, self.p, other.p, self.n * self.dtype.itemsize))
        return self

    def view(self, offset, n):
        """A non-owning sub-array starting at element offset."""
        v = DeviceArray.__new__(DeviceArray)
        v.ctx, v.n, v.dtype = self.ctx, int(n), self.dtype
        v.p = self.p + int(offset) * self.dtype.itemsize
        v._view = True
        return v

    def free(self):
        if getattr(self, "_view", False):
            return
        if self.p and self.ctx.h:
            self.ctx._L.mdg_dev_free(self.ctx.h, self.p)
        self.p = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One grid on the GPU (mdg_create / mdg_destroy)."""

    def __init__(self, Nx, Nz, xmin, xmax, zmin, zmax):
        self._L = lib()
        self.Nx, self.Nz = Nx, Nz
        h = C.c_void_p()
        rc = self._L.mdg_create(C.byref(GridDesc(Nx, Nz, xmin, xmax, zmin, zmax, 0)), C.byref(h))
        if rc != 0:
            raise MdgError(rc, "mdg_create failed (is a GPU visible?)")
        self.h = h
        self.n_markers = 0

    def device(self, n, dtype=np.float64, host=None):
        """Allocate a DeviceArray of n elements (optionally filled from host)."""
        return DeviceArray(self, n, dtype, host)

    def close(self):
        if getattr(self, "h", None):
            self._L.mdg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MdgError(rc, self._L.mdg_last_error(self.h).decode())

    def node_count(self, centroid):
        return {1: (self.Nx - 1) * (self.Nz - 1), 0: self.Nx * self.Nz,
                -1: self.Nx * (self.Nz + 1), -2: (self.Nx + 1) * self.Nz}[centroid]

    def coords(self):
        Nx, Nz = self.Nx, self.Nz
        out = dict(xg=np.zeros(Nx), zg=np.zeros(Nz), xc=np.zeros(Nx - 1), zc=np.zeros(Nz - 1),
                   xvz=np.zeros(Nx + 1), zvx=np.zeros(Nz + 1))
        self._check(self._L.mdg_grid_coords(self.h, *[ptr(out[k]) for k in
                                                      ("xg", "zg", "xc", "zc", "xvz", "zvx")]))
        return out

    def set_bc(self, BCu_t, BCu_v, BCv_t, BCv_v, BCp_t, BCg_t):
        nu, nv = self.node_count(-1), self.node_count(-2)
        self._check(self._L.mdg_set_bc(self.h, checked(BCu_t, np.int8, nu, "BCu_t"),
                                       checked(BCu_v, np.float64, nu, "BCu_v"),
                                       checked(BCv_t, np.int8, nv, "BCv_t"),
                                       checked(BCv_v, np.float64, nv, "BCv_v"),
                                       checked(BCp_t, np.int8, self.node_count(1), "BCp_t"),
                                       checked(BCg_t, np.int8, self.node_count(0), "BCg_t")))

    def markers_set(self, x, z, phase):
        n = len(x)
        self._check(self._L.mdg_markers_set(self.h, n, checked(x, np.float64, n, "x"),
                                            checked(z, np.float64, n, "z"),
                                            checked(phase, np.int32, n, "phase")))
        self.n_markers = n

    def _marker_args(self, x, z, phase):
        n = len(x)
        return n, (checked(x, np.float64, n, "x"), checked(z, np.float64, n, "z"),
                   checked(phase, np.int32, n, "phase"))

    def marker_index(self, centroid):
        ip = np.zeros(self.n_markers, np.int32)
        jp = np.zeros(self.n_markers, np.int32)
        self._check(self._L.mdg_marker_index(self.h, centroid, ptr(ip), ptr(jp)))
        return ip, jp

    def v2p(self, u, v, x, z, phase, conserv=0, vx=None, vz=None):
        """V2P (ParticleRoutines.c:63): (vx, vz) at the markers."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        vx = np.zeros(n) if vx is None else vx
        vz = np.zeros(n) if vz is None else vz
        self._check(self._L.mdg_v2p(self.h, checked(u, np.float64, self.node_count(-1), "u"),
                                    checked(v, np.float64, self.node_count(-2), "v"), conserv, n,
                                    px, pz, pp, checked(vx, np.float64, n, "vx"),
                                    checked(vz, np.float64, n, "vz")))
        return vx, vz

    def advect_rk4(self, u, v, dt, x, z, phase, conserv=0, vx=None, vz=None):
        """RogerGuntherII (AdvectionRoutines.c:347): x, z, phase (and vx, vz
        when given) updated in place; numpy arrays must be contiguous float64
        / int32, or DeviceArray."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        self._check(self._L.mdg_advect_rk4(
            self.h, checked(u, np.float64, self.node_count(-1), "u"),
            checked(v, np.float64, self.node_count(-2), "v"), dt, conserv, n, px, pz, pp,
            checked(vx, np.float64, n, "vx", optional=True),
            checked(vz, np.float64, n, "vz", optional=True)))

    def g2p_centroids(self, F, x, z, phase, out=None):
        """Interp_Grid2P_centroids2 (ParticleRoutines.c:1332): a centre field
        at the markers."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        out = np.zeros(n) if out is None else out
        self._check(self._L.mdg_g2p_centroids(self.h, checked(F, np.float64, self.node_count(1), "F"), n,
                                              px, pz, pp, checked(out, np.float64, n, "out")))
        return out

    def g2p_vertices(self, F, x, z, phase, out=None):
        """Interp_Grid2P (ParticleRoutines.c:1223) on the vertices."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        out = np.zeros(n) if out is None else out
        self._check(self._L.mdg_g2p_vertices(self.h, checked(F, np.float64, self.node_count(0), "F"), n,
                                             px, pz, pp, checked(out, np.float64, n, "out")))
        return out

    def update_particle_stress(self, sxxd, szzd, sxz, sxxd0, szzd0, sxz0, x, z, phase, msxxd, mszzd, msxz):
        """UpdateParticleStress (RheologyParticles.c:1725; no subgrid stress
        diffusion, no rotation): marker stresses updated in place."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        nc, nv = self.node_count(1), self.node_count(0)
        self._check(self._L.mdg_update_particle_stress(
            self.h, checked(sxxd, np.float64, nc, "sxxd"), checked(szzd, np.float64, nc, "szzd"),
            checked(sxz, np.float64, nv, "sxz"), checked(sxxd0, np.float64, nc, "sxxd0"),
            checked(szzd0, np.float64, nc, "szzd0"), checked(sxz0, np.float64, nv, "sxz0"), n, px, pz, pp,
            checked(msxxd, np.float64, n, "msxxd"), checked(mszzd, np.float64, n, "mszzd"),
            checked(msxz, np.float64, n, "msxz")))

    def update_particle_pressure(self, p_in, p0_n, x, z, phase, P, dp=None):
        """UpdateParticlePressure (RheologyParticles.c:1621, no subgrid
        diffusion): P updated in place; returns dp when given."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        nc = self.node_count(1)
        self._check(self._L.mdg_update_particle_pressure(
            self.h, checked(p_in, np.float64, nc, "p_in"), checked(p0_n, np.float64, nc, "p0_n"),
            checked(dp, np.float64, nc, "dp", optional=True), n, px, pz, pp, checked(P, np.float64, n, "P")))
        return dp

    STRAIN_RATES = ("eII_pl", "eII_pwl", "eII_exp", "eII_lin", "eII_gbs", "eII_el", "eII_cst")
    STRAINS = ("strain", "strain_el", "strain_pl", "strain_pl_vol", "strain_pwl", "strain_exp", "strain_lin",
               "strain_gbs")

    def accumulated_strain(self, dt, x, z, phase, rates, strains, div_u_pl=None):
        """AccumulatedStrainII (RheologyParticles.c:346): rates maps names of
        STRAIN_RATES to centre arrays (missing = 0), strains maps names of
        STRAINS to marker arrays updated in place (missing = skipped)."""
        n, (px, pz, pp) = self._marker_args(x, z, phase)
        nc = self.node_count(1)
        e = (C.c_void_p * 7)(*[checked(rates.get(k), np.float64, nc, k, optional=True) for k in self.STRAIN_RATES])
        st = (C.c_void_p * 8)(*[checked(strains.get(k), np.float64, n, k, optional=True) for k in self.STRAINS])
        self._check(self._L.mdg_accumulated_strain(self.h, dt, C.cast(e, C.c_void_p),
                                                   checked(div_u_pl, np.float64, nc, "div_u_pl", optional=True), n,
                                                   px, pz, pp, C.cast(st, C.c_void_p)))

    def rotate_particle_stress(self, mode, dt, phase, wxzm, sxxd, szzd, sxz):
        """Stress rotation of UpdateParticleStress (RheologyParticles.c:1965):
        marker stresses updated in place."""
        n = len(phase)
        self._check(self._L.mdg_rotate_particle_stress(
            self.h, int(mode), dt, n, checked(phase, np.int32, n, "phase"), checked(wxzm, np.float64, n, "wxzm"),
            checked(sxxd, np.float64, n, "sxxd"), checked(szzd, np.float64, n, "szzd"),
            checked(sxz, np.float64, n, "sxz")))

    def update_particle_stress_subgrid(self, dt, G_phase, sxxd, szzd, sxz, sxxd0, szzd0, sxz0, eta_phys_s, x, z,
                                       phase, msxxd, mszzd, msxz, interp_stencil=1):
        """UpdateParticleStress, subgrid_diffusion = 2 (RheologyParticles.c:
        1811-1895): marker stresses updated in place."""
        n, nc, nv = len(phase), (self.Nx - 1) * (self.Nz - 1), self.Nx * self.Nz
        G = np.ascontiguousarray(G_phase, np.float64)
        f = lambda a, m, name: checked(a, np.float64, m, name)
        self._check(self._L.mdg_update_particle_stress_subgrid(
            self.h, dt, ptr(G), G.size, int(interp_stencil), f(sxxd, nc, "sxxd"), f(szzd, nc, "szzd"),
            f(sxz, nv, "sxz"), f(sxxd0, nc, "sxxd0"), f(szzd0, nc, "szzd0"), f(sxz0, nv, "sxz0"),
            f(eta_phys_s, nv, "eta_phys_s"), n, f(x, n, "x"), f(z, n, "z"), checked(phase, np.int32, n, "phase"),
            f(msxxd, n, "msxxd"), f(mszzd, n, "mszzd"), f(msxz, n, "msxz")))

    def update_particle_pressure_subgrid(self, dt, beta_phase, p_in, p0_n, eta_phys_n, x, z, phase, P, dp=None):
        """UpdateParticlePressure, subgrid_diffusion >= 2 (RheologyParticles.c:
        1621-1695): P updated in place; returns dp."""
        n, nc = len(phase), (self.Nx - 1) * (self.Nz - 1)
        B = np.ascontiguousarray(beta_phase, np.float64)
        dp = np.zeros(nc) if dp is None else dp
        f = lambda a, m, name: checked(a, np.float64, m, name)
        self._check(self._L.mdg_update_particle_pressure_subgrid(
            self.h, dt, ptr(B), B.size, f(p_in, nc, "p_in"), f(p0_n, nc, "p0_n"), f(eta_phys_n, nc, "eta_phys_n"),
            f(dp, nc, "dp"), n, f(x, n, "x"), f(z, n, "z"), checked(phase, np.int32, n, "phase"), f(P, n, "P")))
        return dp

    def count_part_cell(self, x, z, phase, n=None, min_part_cell=16, cell_tag=None):
        """Re-seeding of CountPartCell_OLD (ParticleRoutines.c:3224): x, z,
        phase (arrays of the marker capacity, the first n in use) are
        updated in place; cell_tag is BCt.type on the cells (30 = air) or
        None.  Returns (n_out, new_slot, new_src)."""
        cap = len(x)
        n = cap if n is None else int(n)
        nq = 4 * (self.Nx - 1) * (self.Nz - 1)
        slot, src = np.zeros(nq, np.int32), np.zeros(nq, np.int32)
        nn, no = C.c_int(0), C.c_int(0)
        self._check(self._L.mdg_count_part_cell(
            self.h, int(min_part_cell), checked(cell_tag, np.int8, nq // 4, "cell_tag", optional=True), n, cap,
            checked(x, np.float64, cap, "x"), checked(z, np.float64, cap, "z"),
            checked(phase, np.int32, cap, "phase"), ptr(slot), ptr(src), C.byref(nn), C.byref(no)))
        return no.value, slot[:nn.value], src[:nn.value]

    def energy_step(self, dt, rho_n, Cp, T, kx, kz, BCT_t, BCT_v, Qr=None, Wdiss=None, Ha=None, max_its=1000,
                    rel_tol=1e-8):
        """EnergyDirectSolve with SolveThermalPCG (ThermalRoutines.c:74,
        ThermalSolver.c:338): T (cells) updated in place; returns the PCG's
        return value (iterations, or 0 / -1 / -2 / -3)."""
        nc = (self.Nx - 1) * (self.Nz - 1)
        nbc = (self.Nx + 1) * (self.Nz + 1)
        its = C.c_int(0)
        f = lambda a, n, name: checked(a, np.float64, n, name)
        o = lambda a, name: checked(a, np.float64, nc, name, optional=True)
        self._check(self._L.mdg_energy_step(
            self.h, dt, f(rho_n, nc, "rho_n"), f(Cp, nc, "Cp"), o(Qr, "Qr"), o(Wdiss, "Wdiss"), o(Ha, "Ha"), f(kx, self.Nx * (self.Nz + 1), "kx"), f(kz, (self.Nx + 1) * self.Nz, "kz"),
            checked(BCT_t, np.int8, nbc, "BCT_t"), f(BCT_v, nbc, "BCT_v"), int(max_its), rel_tol,
            checked(T, np.float64, nc, "T"), C.byref(its)))
        return its.value

    def markers_gather(self, src, dst_sorted):
        self._check(self._L.mdg_markers_gather(self.h, ptr(src), ptr(dst_sorted)))

    def p2mastah(self, mat_prop, node_field=None, node_type=None, flag=0, avg=0, prop=0,
                 centroid=1, interp_stencil=1, nb_phases=0, phase_perc=None):
        """P2Mastah (ParticleRoutines.c:2170).  Returns node_field, or the
        phase fraction arrays when prop == 1."""
        nn = self.node_count(centroid)
        if prop == 1:
            if phase_perc is None:
                phase_perc = [np.zeros(nn) for _ in range(nb_phases)]
            arr = (C.c_void_p * nb_phases)(*[ptr(a) for a in phase_perc])
            self._check(self._L.mdg_p2mastah(self.h, ptr(mat_prop), 0, None, ptr(node_type),
                                             flag, avg, 1, centroid, interp_stencil, nb_phases,
                                             C.cast(arr, C.c_void_p)))
            return phase_perc
        if node_field is None:
            node_field = np.zeros(nn)
        self._check(self._L.mdg_p2mastah(self.h, checked(mat_prop, np.float64, len(mat_prop), "mat_prop"),
                                         len(mat_prop), checked(node_field, np.float64, nn, "node_field"),
                                         checked(node_type, np.int8, nn, "node_type", optional=True),
                                         flag, avg, 0, centroid, interp_stencil, 0, None))
        return node_field

    def p2mastah_fused(self, centroid, interp_stencil, fields):
        """P2Mastah_Fused; fields = list of dicts(src, dst, flag, avg, node_type)."""
        arr = (P2GField * len(fields))()
        for i, f in enumerate(fields):
            arr[i] = P2GField(ptr(f["src"]), len(f["src"]), ptr(f["dst"]),
                              ptr(f.get("node_type")), f.get("flag", 0), f.get("avg", 0))
        self._check(self._L.mdg_p2mastah_fused(self.h, centroid, interp_stencil, arr, len(fields)))

    # ------------------------------------------------------------- Stokes
    def eqn_count(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        self._check(self._L.mdg_eqn_count(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def stokes_apply(self, D11, D22, D33, x):
        """y = K x (Picard operator, equation numbering)."""
        y = np.zeros(len(x))
        self._check(self._L.mdg_stokes_apply(self.h, ptr(D11), ptr(D22), ptr(D33),
                                             ptr(np.ascontiguousarray(x)), ptr(y)))
        return y

    def apply_bc(self, u, v):
        self._check(self._L.mdg_apply_bc(self.h, ptr(u), ptr(v)))

    def extract_solution(self, sol, alpha, u, v, p):
        """ExtractSolutions2 (StokesRoutines.c:536-576) + ApplyBC, in place."""
        self._check(self._L.mdg_extract_solution(self.h, ptr(sol), float(alpha), ptr(u), ptr(v), ptr(p)))

    def strain_rates(self, u, v, out=None):
        """StrainRateComponents' cell rates and vertex exz; out (optional)
        holds div, exxd, ezzd, exz arrays to write into (e.g. DeviceArray)."""
        nc = (self.Nx - 1) * (self.Nz - 1)
        if out is None:
            out = dict(div=np.zeros(nc), exxd=np.zeros(nc), ezzd=np.zeros(nc),
                       exz=np.zeros(self.Nx * self.Nz))
        self._check(self._L.mdg_strain_rates(self.h, ptr(u), ptr(v), ptr(out["div"]),
                                             ptr(out["exxd"]), ptr(out["ezzd"]), ptr(out["exz"])))
        return out

    def stokes_diag(self, D11, D22, D33, d=None):
        """ExtractDiagonalScale (StokesRoutines.c:1260-1300) scaling vector."""
        if d is None:
            d = np.zeros(self.eqn_count()[0])
        self._check(self._L.mdg_stokes_diag(self.h, ptr(D11), ptr(D22), ptr(D33), ptr(d)))
        return d

    # ---------------------------------------------------------- learned PC
    def pcnet_param_count(self):
        n = C.c_longlong()
        self._check(self._L.mdg_pcnet_param_count(self.h, C.byref(n)))
        return n.value

    def pcnet_set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        self._check(self._L.mdg_pcnet_set_params(self.h, ptr(flat), flat.size))

    def pcnet_get_params(self, grads=False):
        n = self.pcnet_param_count()
        p, g = np.zeros(n, np.float32), np.zeros(n, np.float32)
        self._check(self._L.mdg_pcnet_get_params(self.h, ptr(p), ptr(g)))
        return (p, g) if grads else p

    def pcnet_forward(self, x):
        H, W = x.shape[:2]
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((H, W, 4), np.float32)
        self._check(self._L.mdg_pcnet_forward(self.h, H, W, ptr(x), ptr(y)))
        return y

    def pcnet_grad(self, x, t, mask, accumulate=False):
        H, W = x.shape[:2]
        arrs = [np.ascontiguousarray(a, np.float32) for a in (x, t, mask)]
        loss = C.c_double()
        self._check(self._L.mdg_pcnet_grad(self.h, H, W, *[ptr(a) for a in arrs],
                                           int(accumulate), C.byref(loss)))
        return loss.value

    def pcnet_adam(self, lr):
        self._check(self._L.mdg_pcnet_adam(self.h, C.c_float(lr)))

    def pcnet_capture(self, max_pairs):
        self._check(self._L.mdg_pcnet_capture(self.h, int(max_pairs)))

    def pcnet_capture_outer(self, max_pairs):
        self._check(self._L.mdg_pcnet_capture_outer(self.h, int(max_pairs)))

    def pcnet_captured(self):
        n, H, W = C.c_int(), C.c_int(), C.c_int()
        self._check(self._L.mdg_pcnet_captured(self.h, C.byref(n), C.byref(H), C.byref(W)))
        return n.value, H.value, W.value

    def pcnet_get_pair(self, i):
        _, H, W = self.pcnet_captured()
        x, t, m = (np.zeros((H, W, 4), np.float32) for _ in range(3))
        self._check(self._L.mdg_pcnet_get_pair(self.h, int(i), ptr(x), ptr(t), ptr(m)))
        return x, t, m

    def pcnet_train(self, steps, batch=1, lr=1e-3, seed=0):
        losses = np.zeros(max(steps, 1))
        self._check(self._L.mdg_pcnet_train(self.h, int(steps), int(batch), C.c_float(lr),
                                            C.c_ulonglong(seed), ptr(losses)))
        return losses[:steps]

    def pcnet_train_residual(self, steps, batch=1, lr=1e-3, seed=0):
        losses = np.zeros(max(steps, 1))
        self._check(self._L.mdg_pcnet_train_residual(self.h, int(steps), int(batch), C.c_float(lr),
                                                     C.c_ulonglong(seed), ptr(losses)))
        return losses[:steps]

    def c2v(self, cent, vert=None, periodic_x=0):
        """InterpCentroidsToVerticesDouble (GridRoutines.c:108-212)."""
        if vert is None:
            vert = np.zeros(self.Nx * self.Nz)
        self._check(self._L.mdg_c2v(self.h, int(periodic_x), ptr(cent), ptr(vert)))
        return vert

    def strain_rates_nodes(self, exxd, ezzd, exz, div, out=None):
        """exz_n, exxd_s, ezzd_s, div_s of StrainRateComponents
        (RheologyDensity.c:2066-2094); out as for strain_rates."""
        nc, nv = (self.Nx - 1) * (self.Nz - 1), self.Nx * self.Nz
        if out is None:
            out = dict(exz_n=np.zeros(nc), exxd_s=np.zeros(nv), ezzd_s=np.zeros(nv),
                       div_s=np.zeros(nv))
        self._check(self._L.mdg_strain_rates_nodes(
            self.h, ptr(exxd), ptr(ezzd), ptr(exz), ptr(div), ptr(out["exz_n"]),
            ptr(out["exxd_s"]), ptr(out["ezzd_s"]), ptr(out["div_s"])))
        return out

    def material_grid(self, phases, centroid, perc, strain_pl=None, style=0):
        """ShearModCompExpGrid + CohesionFrictionDilationGrid on one node set."""
        n = self.node_count(centroid)
        perc = np.ascontiguousarray(perc, np.float64)
        out = {k: np.zeros(n) for k in ("mu", "bet", "C", "fric", "dil")}
        arr = (PhaseSoft * len(phases))(*phases)
        sp = None if strain_pl is None else np.ascontiguousarray(strain_pl, np.float64)
        self._check(self._L.mdg_material_grid(self.h, arr, len(phases), int(centroid), int(style),
                                              ptr(perc), ptr(sp), *[ptr(out[k]) for k in
                                                                    ("mu", "bet", "C", "fric",
                                                                     "dil")]))
        return out

    @staticmethod
    def _rheo_inputs(arrays, n, nph):
        """mdg_rheo_nodes with the inputs of arrays (numpy arrays converted
        to contiguous float64 and checked to hold n values, nph*n for perc;
        DeviceArrays passed as they are)."""
        f, keep = RheoNodes(), []
        for k in RHEO_IN:
            a = arrays.get(k)
            if isinstance(a, np.ndarray):
                a = np.ascontiguousarray(a, np.float64)
                want = n * nph if k == "perc" else n
                if a.size != want:
                    raise ValueError("%s holds %d values, expected %d" % (k, a.size, want))
                keep.append(a)
            setattr(f, k, ptr(a))
        return f, keep

    def tangent_grid(self, phases, prm, centroid, D=None, **arrays):
        """Newton tangent coefficients: (8, cells) or (4, vertices); D may be
        a DeviceArray of that many values to write into."""
        n = self.node_count(centroid)
        if D is None:
            D = np.zeros(((8 if centroid else 4), n))
        f, keep = self._rheo_inputs(arrays, n, len(phases))
        ph = (PhaseProp * len(phases))(*phases)
        self._check(self._L.mdg_tangent_grid(self.h, ph, len(phases), C.byref(prm), int(centroid),
                                             C.byref(f), ptr(D)))
        return D

    def jacobian_apply(self, Dn, Ds, x):
        y = np.zeros(len(x))
        self._check(self._L.mdg_jacobian_apply(self.h, ptr(np.ascontiguousarray(Dn)),
                                               ptr(np.ascontiguousarray(Ds)),
                                               ptr(np.ascontiguousarray(x)), ptr(y)))
        return y

    def newton_solver(self, Dn, Ds, D11, D33, rhs_mom, rhs_cont, accept_diverged=False, sol=None, **opts):
        """mdg_newton_solver: returns (sol, SolverOpts); the arrays may be
        DeviceArray (sol then receives the solution on the device)."""
        o = SolverOpts()
        for k, v in opts.items():
            setattr(o, k, v)
        if sol is None:
            sol = np.zeros(len(rhs_mom) + len(rhs_cont))
        rc = self._L.mdg_newton_solver(self.h, C.byref(o), ptr(_flat(Dn)), ptr(_flat(Ds)), ptr(D11), ptr(D33),
                                       ptr(_flat(rhs_mom)), ptr(_flat(rhs_cont)), ptr(sol))
        self._solved(rc, o, accept_diverged)
        return sol, o

    def viscosity_grid(self, phases, prm, centroid, outputs=RHEO_OUT, **arrays):
        """NonNewtonianViscosityGrid (RheologyDensity.c:1131-1426) on the
        centres (centroid=1) or vertices (centroid=0).  arrays holds the
        inputs named in RHEO_IN (perc has shape (nph, n)); outputs named in
        RHEO_OUT are allocated on the host and returned in a dict, or may be
        passed in arrays (e.g. as DeviceArray)."""
        n = self.node_count(centroid)
        out = {}
        f, keep = self._rheo_inputs(arrays, n, len(phases))
        for k in RHEO_OUT:
            a = arrays.get(k)
            if a is None and k in outputs:
                a = np.zeros(n)
            if a is not None:
                out[k] = a
            setattr(f, k, ptr(a))
        ph = (PhaseProp * len(phases))(*phases)
        self._check(self._L.mdg_viscosity_grid(self.h, ph, len(phases), C.byref(prm),
                                               int(centroid), C.byref(f)))
        return out

    def eval_rhs(self, gx, gz, rho_n, rho_s, rx=None, rz=None):
        if rx is None:
            rx = np.zeros(self.Nx * (self.Nz + 1))
        if rz is None:
            rz = np.zeros((self.Nx + 1) * self.Nz)
        self._check(self._L.mdg_eval_rhs(self.h, gx, gz, ptr(rho_n), ptr(rho_s), ptr(rx), ptr(rz)))
        return rx, rz

    def stokes_residual(self, sxxd, szzd, sxz, p_corr, div, roger_x, roger_z, d=None, F=None):
        neq = self.eqn_count()[0]
        if F is None:
            F = np.zeros(neq)
        norms = np.zeros(3)
        self._check(self._L.mdg_stokes_residual(self.h, ptr(sxxd), ptr(szzd), ptr(sxz),
                                                ptr(p_corr), ptr(div), ptr(roger_x),
                                                ptr(roger_z), ptr(d), ptr(F), ptr(norms)))
        return F, norms

    def stokes_residual_static(self, sxxd, szzd, sxz, p_corr, div, roger_x, roger_z, F, d=None):
        """mdg_stokes_residual_static: the residual without any allocation;
        every array a DeviceArray (F receives the residual), returns norms."""
        arrs = (sxxd, szzd, sxz, p_corr, div, roger_x, roger_z, F) + ((d,) if d is not None else ())
        if not all(isinstance(a, DeviceArray) for a in arrs):
            raise TypeError("stokes_residual_static takes DeviceArray arguments only")
        norms = np.zeros(3)
        self._check(self._L.mdg_stokes_residual_static(self.h, ptr(sxxd), ptr(szzd), ptr(sxz), ptr(p_corr),
                                                        ptr(div), ptr(roger_x), ptr(roger_z), ptr(d), ptr(F),
                                                        ptr(norms)))
        return norms

    def _solved(self, rc, o, accept_diverged):
        # MDG_ERR_DIVERGED leaves the last iterate in sol; the reference's
        # KillerSolver carries on with it after max_its (Solvers.c:1298-1334)
        if rc == MDG_ERR_DIVERGED and accept_diverged:
            o.status = rc
            return
        self._check(rc)

    def killer_solver(self, D11, D22, D33, rhs_mom, rhs_cont, accept_diverged=False, sol=None, **opts):
        """Replacement of KillerSolver: returns (sol, SolverOpts).  With
        accept_diverged, a solve that misses its tolerance returns its last
        iterate with o.status = MDG_ERR_DIVERGED instead of raising.  The
        arrays may be DeviceArray (sol then receives the solution there)."""
        o = SolverOpts()
        for k, v in opts.items():
            setattr(o, k, v)
        if sol is None:
            sol = np.zeros(len(rhs_mom) + len(rhs_cont))
        rc = self._L.mdg_killer_solver(self.h, C.byref(o), ptr(D11), ptr(D22), ptr(D33),
                                       ptr(_flat(rhs_mom)), ptr(_flat(rhs_cont)), ptr(sol))
        self._solved(rc, o, accept_diverged)
        return sol, o

    def scale_into(self, x, a, y):
        self._check(self._L.mdg_dev_scale(self.h, len(x), a, ptr(x), ptr(y)))

    def killer_solver_dev(self, D11, D22, D33, F, nm, sol, **opts):
        """killer_solver on device arrays: F and sol hold [mom; cont]."""
        o = SolverOpts()
        for k, v in opts.items():
            setattr(o, k, v)
        rc = self._L.mdg_killer_solver(self.h, C.byref(o), ptr(D11), ptr(D22), ptr(D33), ptr(F),
                                       ptr(F.view(nm, len(F) - nm)), ptr(sol))
        self._check(rc)
        return o

    def time_smoother(self, reps=50):
        ms, b = C.c_double(), C.c_double()
        self._check(self._L.mdg_time_smoother(self.h, reps, C.byref(ms), C.byref(b)))
        return ms.value, b.value

    def last_kernel_ms(self):
        v = C.c_double()
        self._check(self._L.mdg_last_kernel_ms(self.h, C.byref(v)))
        return v.value

    def last_weights_ms(self):
        v = C.c_double()
        self._check(self._L.mdg_last_weights_ms(self.h, C.byref(v)))
        return v.value
