"""Reader for MDOODZ's `.txt` model files, for the keys the per-step path
reads: the grid, scales, time stepping, linear / non-linear / thermal solver
settings, marker settings, gravity and the per-phase material parameters
(ReadInputFile, InputOutput.c:1077-1560).  Values come out in the model's
scaled units, exactly as ReadInputFile leaves them in `params`, `Nparams`
and `mat_prop`, and convert to this library's structures (mdg_phase_prop,
mdg_phase_soft, mdg_rheo_params, mdg_solver_opts).

Lookup rules of the reference's readers, kept because input files rely on
them:
* a line's key is its text up to the first space; lines without a space,
  or starting with one, hold no key (ReadDou2 :2130-2200, ReadInt2
  :2053-2124);
* a top-level key takes the first matching line of the whole file
  (phase blocks included) and the number after the first '=' (C atof /
  atoi: leading part only, so trailing comments are ignored);
* a phase key is searched between the line `ID = k` and the next `ID`
  line (ReadMatProps :2207-2350);
* a string key matches a line by prefix and takes the word that starts
  two characters after '=' (ReadChar :1873-1960).
"""
import math
import re

from . import flowlaws as FL

RG = 8.314510      # gas constant of the reference (mdoodz.h)
ZERO_C = 273.15    # 0 degC in K (mdoodz.h)

_FLOAT = re.compile(r"\s*([+-]?(?:(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|inf(?:inity)?|nan))", re.IGNORECASE)
_INT = re.compile(r"\s*([+-]?\d+)")


def c_atof(s):
    """C atof: the longest leading decimal number, 0.0 if there is none."""
    m = _FLOAT.match(s)
    return float(m.group(1)) if m else 0.0


def c_atoi(s):
    """C atoi: the leading integer, 0 if there is none."""
    m = _INT.match(s)
    return int(m.group(1)) if m else 0


def _key(line):
    i = line.find(" ")
    if i <= 0:
        return None
    return line[:i]


class InputFile:
    """The reference's key lookups over the lines of one model file."""

    def __init__(self, text):
        self.lines = text.splitlines(keepends=True)

    @classmethod
    def read(cls, path):
        with open(path, "r", encoding="utf-8", errors="replace") as f:
            return cls(f.read())

    def _raw(self, key):
        for line in self.lines:
            if _key(line) == key and "=" in line:
                return line[line.index("=") + 1:]
        return None

    def dou(self, key, default):
        r = self._raw(key)
        return default if r is None else c_atof(r)

    def int(self, key, default):
        r = self._raw(key)
        return default if r is None else c_atoi(r)

    def char(self, key, default):
        for line in self.lines:
            if line.startswith(key) and "=" in line:
                word = re.match(r"[^\s]*", line[line.index("=") + 2:]).group(0)
                return word
        return default

    def phase_ids(self):
        return [c_atoi(line[line.index("=") + 1:]) for line in self.lines
                if _key(line) == "ID" and "=" in line]

    def mat(self, key, phase, default):
        n = len(self.lines)
        i = 0
        while i < n:
            line = self.lines[i]
            i += 1
            if _key(line) == "ID" and "=" in line and c_atoi(line[line.index("=") + 1:]) == phase:
                while i < n:
                    pl = self.lines[i]
                    i += 1
                    k = _key(pl)
                    if k == "ID":
                        return default
                    if k == key and "=" in pl:
                        return c_atof(pl[pl.index("=") + 1:])
                return default
        raise ValueError("no phase ID %d in the input file" % phase)


# Top-level keys: name -> (kind, default, scale attribute or None).  Scaled
# values are divided by that scale, as ReadInputFile does.
_MODEL_KEYS = {
    # space
    "Nx": ("i", 10, None), "Nz": ("i", 10, None),
    "xmin": ("d", -1.0, "L"), "xmax": ("d", 1.0, "L"), "zmin": ("d", -1.0, "L"), "zmax": ("d", 1.0, "L"),
    # time
    "Nt": ("i", 1, None), "dt": ("d", 0.0, "t"), "t_end": ("d", 0.0, "t"), "Courant": ("d", 0.5, None),
    "RK": ("i", 4, None), "constant_dt": ("i", 0, None), "stress_rotation": ("i", 1, None),
    "dt_max": ("d", 1e20, "t"), "dt_min": ("d", -1e20, "t"),
    # switches
    "mechanical": ("i", 1, None), "advection": ("i", 1, None), "elastic": ("i", 0, None),
    "thermal": ("i", 0, None), "anisotropy": ("i", 0, None), "polar": ("i", 0, None),
    "finite_strain": ("i", 0, None), "compressible": ("i", 0, None), "melting": ("i", 0, None),
    # linear solver
    "penalty": ("d", 1.0e3, None), "diag_scaling": ("i", 1, None), "preconditioner": ("i", 0, None),
    "lin_abs_div": ("d", 1.0e-9, None), "lin_rel_div": ("d", 1.0e-5, None),
    "lin_abs_mom": ("d", 1.0e-9, None), "lin_rel_mom": ("d", 1.0e-5, None),
    "lin_solver": ("i", 2, None), "max_its_PH": ("i", 10, None),
    "max_its_KSP": ("i", 100, None), "rel_tol_KSP": ("d", 1e-4, None),
    # thermal solver
    "thermal_solver": ("i", 0, None), "max_its_thermal": ("i", 1000, None), "rel_tol_thermal": ("d", 1e-8, None),
    # non-linear solver
    "nit_max": ("i", 1, None), "Newton": ("i", 0, None), "Picard2Newton": ("i", 0, None),
    "Picard2Newton_tol": ("d", 1e-1, None), "max_its_Pic": ("i", 10, None), "let_res_grow": ("i", 0, None),
    "nonlin_abs_mom": ("d", 1.0e-6, None), "nonlin_abs_div": ("d", 1.0e-6, None),
    "nonlin_rel_mom": ("d", 1.0e-6, None), "nonlin_rel_div": ("d", 1.0e-6, None),
    "min_eta": ("d", 1e18, "eta"), "max_eta": ("d", 1e24, "eta"), "eta_tol": ("d", 1e11, None),
    "safe_mode": ("i", 0, None), "safe_dt_div": ("d", 5.0, None), "max_num_stag": ("i", 3, None),
    "line_search": ("i", 0, None), "pc_line_search": ("i", 0, None), "line_search_min": ("d", 0.0, None),
    # grid / markers
    "eta_average": ("i", 0, None), "interp_stencil": ("i", 1, None), "interp_mode": ("i", 0, None),
    "subgrid_diffusion": ("i", 0, None), "conserv_interp": ("i", 0, None), "direct_neighbour": ("i", 0, None),
    "reseed_markers": ("i", 1, None), "reseed_mode": ("i", 0, None),
    "Nx_part": ("i", 4, None), "Nz_part": ("i", 4, None), "min_part_cell": ("i", 16, None),
    # boundary conditions and setup
    "shear_style": ("i", 0, None), "periodic_x": ("i", 0, None), "pure_shear_ALE": ("i", 0, None),
    "free_surface": ("i", 0, None), "free_surface_stab": ("d", 0.0, None),
    "shear_heating": ("i", 1, None), "adiab_heating": ("i", 0, None),
    "bkg_strain_rate": ("d", 1e-30, "E"), "bkg_div_rate": ("d", 0.0, "E"), "bkg_pressure": ("d", 0.0, "S"),
    "bkg_temperature": ("d", 0.0, "T"),
    "force_act_vol_ast": ("i", 0, None), "act_vol_dis_ast": ("d", 0.0, None), "act_vol_dif_ast": ("d", 0.0, None),
    "force_melt_weak": ("i", 0, None), "melt_weak": ("d", 0.0, None),
    "gx": ("d", 0.0, "a"), "gz": ("d", 0.0, "a"),
    "Nb_phases": ("i", 0, None),
}
_MODEL_KEYS.update({"user%d" % i: ("d", 0.0, None) for i in range(10)})


class Phase:
    """One phase of mat_prop in scaled units (InputOutput.c:1400-1530)."""

    def __init__(self, f, k, scl, model):
        d = lambda key, default: f.mat(key, k, default)
        self.rho = d("rho", 2700.0) / scl.rho
        self.G = d("G", 1.0e10) / scl.S
        self.Cp = d("Cp", 1.0e3) / scl.Cp
        self.k = d("k", 1.0e-6) / scl.k
        self.Qr = d("Qr", 1.0e-30) / (scl.W / scl.L ** 3)
        self.alp = d("alp", 0.0) / (1.0 / scl.T)
        self.bet = d("bet", 1.0e-40) / (1.0 / scl.S)
        self.drho = d("drho", 0.0) / scl.rho
        self.plast = int(d("plast", 1.0))
        self.C = d("C", 1.0e7) / scl.S
        self.phi = d("phi", 30.0) * math.pi / 180.0
        self.psi = d("psi", 0.0) * math.pi / 180.0
        self.Slim = d("Slim", 1.0e90) / scl.S
        self.n_vp = d("n_vp", 1.0)
        self.eta_vp = d("eta_vp", 0.0) / scl.S / math.pow(scl.t, 1.0 / self.n_vp)
        if self.n_vp > 1.0001:
            raise ValueError("phase %d: power-law viscoplasticity (n_vp > 1) is not implemented" % k)
        self.coh_soft = int(d("coh_soft", 0.0))
        self.phi_soft = int(d("phi_soft", 0.0))
        self.psi_soft = int(d("psi_soft", 0.0))
        self.C_end = d("Ce", self.C * scl.S) / scl.S
        self.phi_end = d("phie", self.phi * 180.0 / math.pi) * math.pi / 180.0
        self.psi_end = d("psie", self.psi * 180.0 / math.pi) * math.pi / 180.0
        self.H_c = d("Hc", 0.0) / scl.S
        self.pls_start = d("plss", 1.0)
        self.pls_end = d("plse", 2.0)
        self.cstv = int(d("cstv", 0.0))
        self.pwlv = int(d("pwlv", 0.0))
        self.linv = int(d("linv", 0.0))
        self.gbsv = int(d("gbsv", 0.0))
        self.expv = int(d("expv", 0.0))
        self.eps0 = d("eps0", 1e-14) / scl.E
        self.tau0 = d("tau0", 2e6) / scl.S
        self.eta0 = d("eta0", 1.0e20) / scl.eta
        self.npwl = d("npwl", 1.0)
        self.Qpwl = d("Qpwl", 0.0)
        self.Apwl = d("Apwl", 0.0)
        self.apwl = d("apwl", 0.0)
        self.tpwl = d("tpwl", 0.0)
        self.pref_pwl = d("pref_pwl", 1.0)
        self.density_model = int(d("density_model", 3))
        self.pwl = None
        if self.pwlv != 0:
            self.pwl = FL.power_law(self.pwlv, scl, n=self.npwl, Q=self.Qpwl, A=self.Apwl, a=self.apwl,
                                    t=self.tpwl, eta0=self.eta0, eps0=self.eps0, tau0=self.tau0,
                                    force_act_vol_ast=model["force_act_vol_ast"],
                                    act_vol_dis_ast=model["act_vol_dis_ast"],
                                    force_melt_weak=model["force_melt_weak"], melt_weak=model["melt_weak"])
        self.gs_ref = d("gs_ref", 2.0e-3) / scl.L
        self.lin = None if self.linv == 0 else FL.linear_law(self.linv, scl, model["force_act_vol_ast"],
                                                              model["act_vol_dif_ast"])
        self.gbs = None if self.gbsv == 0 else FL.gbs_law(self.gbsv, scl)
        self.exp = None if self.expv == 0 else FL.exponential_law(self.expv, scl)


class Model:
    """params + Nparams + mat_prop of one input file, scaled."""

    def __init__(self, f):
        self.file = f
        self.scaling = FL.Scale(eta=f.dou("eta", 1.0), L=f.dou("L", 1.0), V=f.dou("V", 1.0), T=f.dou("T", 1.0))
        s = self.scaling
        p = {}
        for key, (kind, default, sc) in _MODEL_KEYS.items():
            v = f.int(key, default) if kind == "i" else f.dou(key, default)
            p[key] = v / getattr(s, sc) if sc else v
        # consequential settings (InputOutput.c:1325-1345)
        if p["interp_stencil"] not in (1, 9):
            raise ValueError("interp_stencil must be 1 or 9")
        if p["shear_style"] == 1:
            p["periodic_x"] = 1
        if p["shear_style"] == 0:
            p["periodic_x"] = 0
        if p["Newton"] == 0:
            p["Picard2Newton"] = 0
        if p["Newton"] == 1:
            p["line_search"] = 1
        if p["lin_solver"] == 3:
            raise ValueError("lin_solver = 3 has been removed from the reference")
        if p["lin_solver"] == 0 or p["Newton"] == 1 or p["anisotropy"] == 1:
            p["lin_solver"] = 2
        self.params = p
        self.writer_subfolder = f.char("writer_subfolder", "./")
        self.R = RG / (s.J / s.T)
        self.phases = [Phase(f, k, s, p) for k in range(p["Nb_phases"])]

    @classmethod
    def read(cls, path):
        return cls(InputFile.read(path))

    def __getitem__(self, key):
        return self.params[key]

    def phase_props(self):
        """mdg_phase_prop per phase (constant viscosity, power-law, diffusion,
        grain-boundary-sliding and Peierls creep, Drucker-Prager)."""
        from . import PhaseProp
        out = []
        for ph in self.phases:
            kw = dict(cstv=1 if ph.cstv else 0, pwlv=0, plast=ph.plast, eta0=ph.eta0, pref_pwl=1.0, F_pwl=1.0,
                      A_pwl=1.0, n_pwl=1.0, Q_pwl=0.0, V_pwl=0.0, C=ph.C, fric=ph.phi, dil=ph.psi, Slim=ph.Slim,
                      eta_vp=ph.eta_vp, n_vp=ph.n_vp, beta=ph.bet, gs_ref=ph.gs_ref)
            if ph.pwl is not None:
                kw.update(FL.phase_prop_fields(ph.pwl, ph.pref_pwl))
            if ph.lin is not None:
                kw.update(FL.linear_prop_fields(ph.lin))
            if ph.gbs is not None:
                kw.update(FL.gbs_prop_fields(ph.gbs))
            if ph.exp is not None:
                kw.update(FL.exponential_prop_fields(ph.exp))
            out.append(PhaseProp(**kw))
        return out

    def phase_soft(self):
        """mdg_phase_soft per phase."""
        from . import PhaseSoft
        return [PhaseSoft(G=ph.G, beta=ph.bet, C=ph.C, C_end=ph.C_end, fric=ph.phi, fric_end=ph.phi_end, dil=ph.psi,
                          dil_end=ph.psi_end, pls_start=ph.pls_start, pls_end=ph.pls_end, H_c=ph.H_c,
                          fric_soft=ph.phi_soft, dil_soft=ph.psi_soft, coh_soft=ph.coh_soft) for ph in self.phases]

    def rheo_params(self, step=1, dt=None):
        """mdg_rheo_params for time step `step` (ViscosityConcise's constants,
        RheologyDensity.c:393,427,496,506: Peierls creep below 1200 degC,
        Tiimin = 1e5 Pa, T floor 0 degC, smallest grain size 1e-13 m)."""
        from . import RheoParams
        s, p = self.scaling, self.params
        return RheoParams(step=step, elastic=p["elastic"], average=p["eta_average"],
                          dt=p["dt"] if dt is None else dt, R=self.R, zeroC=ZERO_C / s.T, min_eta=p["min_eta"],
                          max_eta=p["max_eta"], eta_tol=p["eta_tol"], Tiimin=1e5 / s.S, scale_E=s.E,
                          scale_eta=s.eta, T_peierls=(1200.0 + ZERO_C) / s.T, d_min=1e-13 / s.L)

    def solver_opts(self):
        """mdg_solver_opts running KillerSolver's own iteration (method 1)
        with the file's penalty and tolerances."""
        from . import SolverOpts
        p = self.params
        return SolverOpts(method=1, penalty=p["penalty"], max_its_PH=p["max_its_PH"], rel_tol_KSP=p["rel_tol_KSP"],
                          max_its_KSP=p["max_its_KSP"], lin_abs_mom=p["lin_abs_mom"], lin_rel_mom=p["lin_rel_mom"],
                          lin_abs_div=p["lin_abs_div"], lin_rel_div=p["lin_rel_div"])

    def loop_options(self):
        """(nit_max, tol, keyword arguments) of PicardLoop.run for the file's
        non-linear settings (Main_DOODZ.c:878-1062 reads the same keys):
        tol = (nonlin_abs_mom, nonlin_rel_mom, nonlin_abs_div, nonlin_rel_div)."""
        p = self.params
        tol = (p["nonlin_abs_mom"], p["nonlin_rel_mom"], p["nonlin_abs_div"], p["nonlin_rel_div"])
        kw = dict(line_search=bool(p["line_search"]), newton=bool(p["Newton"]),
                  picard2newton=bool(p["Picard2Newton"]), picard2newton_tol=p["Picard2Newton_tol"],
                  max_its_pic=p["max_its_Pic"], line_search_min=p["line_search_min"],
                  let_res_grow=p["let_res_grow"], pc_line_search=p["pc_line_search"], safe_mode=p["safe_mode"],
                  safe_dt_div=p["safe_dt_div"], max_num_stag=p["max_num_stag"])
        return p["nit_max"], tol, kw
