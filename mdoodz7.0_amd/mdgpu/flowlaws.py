"""Creep flow-law databases and the model scales.

Restates ReadDataPowerLaw (FlowLaws.c:46-588), ReadDataLinear (:591-706),
ReadDataGBS (:711-776), ReadDataExponential (:782-842) and ScaleMe
(InputOutput.c:1835-1849).  A flow law is the published laboratory fit

    strain rate = A * stress^n * d^-m * f^r * exp(a phi) * exp(-(Q + P V) / (R T)),

stored with its experiment type t (0: already in invariant form,
1: axial compression, 2: simple shear).  ViscosityConcise
(RheologyDensity.c:502) turns it into the invariant-form viscosity
eta = B Eii^(1/n - 1) with

    B = pref * F * A^(-1/n) * exp((Q + P V) / (n R T)) * d^(m/n) * f^(-r/n) * exp(-a phi / n),

where F converts the experiment's stress and strain-rate measures to the
second invariants.  All values leave this module in the model's scaled
units, like mat_prop after ReadInputFile.

Diffusion (linear) creep and grain-boundary sliding use the same form and
correction factor; Peierls creep is the exponential law

    strain rate = E * exp(-Q / (R T) * (1 - sigma / S)^2) * sigma^n,

regularised in ViscosityConcise (RheologyDensity.c:516-523) by its tangent
power law at sigma = G * S.

The tables hold the published constants each case of the readers sets
(same numbers, same case numbers, so input files keep their meaning).
"""
import math
from dataclasses import dataclass

# number: (label, t, n, m, r, Q [J/mol], V [m^3/mol], A [Pa^-n s^-1], f [Pa], a)
POWER_LAWS = {
    10: ("Westerly granite (dry), Hansen & Carter 1983", 1, 3.3, 0.0, 0.0, 186.5e3, 0.0, 3.1623e-26, 0.0, 0.0),
    11: ("Maryland diabase, Mackwell et al. 1998", 1, 4.7, 0.0, 0.0, 485.0e3, 0.0, 5.0477e-28, 0.0, 0.0),
    12: ("Maryland diabase, Carter & Tsenn 1987", 1, 3.0, 0.0, 0.0, 276.0e3, 0.0, 3.2e-20, 0.0, 0.0),
    13: ("Wet quartzite, Ranalli 1995", 1, 2.3, 0.0, 0.0, 154.0e3, 0.0, 5.0717e-18, 0.0, 0.0),
    14: ("Rock salt, Ranalli 1995", 1, 5.3, 0.0, 0.0, 102.0e3, 0.0, 9.9848e-32, 0.0, 0.0),
    15: ("Calcite, Renner et al. 2002", 1, 4.7, 0.0, 0.0, 297.0e3, 0.0, 1.5849e-25, 0.0, 0.0),
    16: ("Felsic granulite, Ranalli 1995", 1, 3.1, 0.0, 0.0, 243.0e3, 0.0, 2.0095e-21, 0.0, 0.0),
    17: ("Mafic granulite, Ranalli 1995", 1, 4.2, 0.0, 0.0, 445.0e3, 0.0, 8.8334e-22, 0.0, 0.0),
    18: ("Calcite, Schmid et al. 1977", 1, 4.7, 0.0, 0.0, 297.0e3, 0.0, 1.5849e-25, 0.0, 0.0),
    19: ("Anorthite 100 (0.007 wt% H2O), Rybacki & Dresen 2000, 2004", 1, 3.0, 0.0, 0.0, 356.0e3, 0.0, 3.9811e-16,
         0.0, 0.0),
    20: ("Garnet, Ji & Martignole 1994", 1, 2.22, 0.0, 0.0, 485.0e3, 0.0, 2.7952e-07, 0.0, 0.0),
    21: ("Omphacite, Zhang et al. 2006", 1, 3.5, 0.0, 0.0, 310.0e3, 0.0, 1.0e-23, 0.0, 0.0),
    22: ("Quartz, Koch et al. 1989", 1, 2.72, 0.0, 0.0, 134.0e3, 0.0, 7.3191e-24, 0.0, 0.0),
    23: ("Feldspar, Shelton & Tullis 1981", 1, 3.2, 0.0, 0.0, 238.0e3, 0.0, 2.0632e-23, 0.0, 0.0),
    24: ("Mica, Kronenberg et al. 1990", 1, 18.0, 0.0, 0.0, 51.0e3, 0.0, 1.0e-138, 0.0, 0.0),
    25: ("Plagioclase, Shelton 1981", 1, 4.0, 0.0, 0.0, 431.0e3, 0.0, 7.9433e-24, 0.0, 0.0),
    26: ("Wet quartzite, Burov 2011 and sources", 1, 2.4, 0.0, 0.0, 160.0e3, 0.0, 3.981e-19, 0.0, 0.0),
    27: ("Dry quartz, Ranalli 1995, 1997, 2003", 1, 2.4, 0.0, 0.0, 156.0e3, 0.0, 2.667e-20, 0.0, 0.0),
    28: ("Quartz, Hirth et al. 2001 (uniaxial)", 1, 4.0, 0.0, 0.0, 135.0e3, 0.0, 6.3096e-36, 0.0, 0.0),
    29: ("Basalt, Hacker & Christie 1990", 0, 3.7, 0.0, 0.0, 244e3, 0.0, 6.9405e-27, 0.0, 0.0),
    30: ("Anorthite (dry), Rybacki & Dresen 2000", 1, 3.0, 0.0, 0.0, 648.0e3, 0.0, 5.0119e-06, 0.0, 0.0),
    31: ("Clinopyroxene (dry), Bystricky & Mackwell 2001", 1, 4.7, 0.0, 0.0, 760.0e3, 0.0, 3.9811e-19, 0.0, 0.0),
    32: ("Anorthite 60, Rybacki & Dresen 2004", 1, 3.0, 0.0, 0.0, 235.0e3, 0.0, 3.16228e-20, 0.0, 0.0),
    33: ("Garnet with activation volume, Ji & Martignole 1994", 1, 2.22, 0.0, 0.0, 485.0e3, 10e-6, 2.7952e-07, 0.0,
         0.0),
    34: ("Westerly granite (dry) with melt weakening, Hansen & Carter 1983", 1, 3.3, 0.0, 0.0, 186.5e3, 0.0,
         3.1623e-26, 0.0, 50.0),
    40: ("Olivine dry dislocation creep, Hirth & Kohlstedt 2003", 1, 3.5, 0.0, 0.0, 530.0e3, 11.0e-6, 1.1000e-16,
         0.0, 0.0),
    41: ("Olivine wet dislocation creep, Hirth & Kohlstedt 2003", 1, 3.5, 0.0, 1.2, 480.0e3, 11.0e-6, 5.6786e-27,
         1.0e9, 0.0),
    42: ("Dry olivine average, Ranalli 1995 / Kirby & Kronenberg 1987", 1, 3.5, 0.0, 0.0, 532.0e3, 0.0, 2.5119e-17,
         0.0, 0.0),
    43: ("Wet olivine average, Ranalli 1995 / Kirby & Kronenberg 1987", 1, 4.0, 0.0, 0.0, 471.0e3, 0.0, 1.9953e-21,
         0.0, 0.0),
    44: ("Olivine, Chopra & Paterson 1984", 1, 3.0, 0.0, 0.0, 520.0e3, 0.0, 1.0e-14, 0.0, 0.0),
    45: ("Olivine dry dislocation creep (Q = 520 kJ), Hirth & Kohlstedt 2003", 1, 3.5, 0.0, 0.0, 520.0e3, 0.0, 1.60e-18,
         0.0, 0.0),
    46: ("Serpentine 1 GPa, Hilairet et al. 2007", 1, 5.8, 0.0, 0.0, 17.6e3, 0.0, 5.3443e-41, 0.0, 0.0),
}


@dataclass
class Scale:
    """The four characteristic scales of an input file (SCALES block) and
    the derived ones of ScaleMe (InputOutput.c:1835-1849)."""
    eta: float
    L: float
    V: float
    T: float

    def __post_init__(self):
        self.t = self.L / self.V
        self.a = self.V / self.t
        self.E = 1.0 / self.t
        self.S = self.eta / self.t
        self.m = self.S * self.L * self.t ** 2
        self.rho = self.m / self.L ** 3
        self.F = self.m * self.L / self.t ** 2
        self.J = self.m * self.L ** 2 / self.t ** 2
        self.W = self.J / self.t
        self.Cp = self.J / self.m / self.T
        self.k = self.W / self.L / self.T
        self.rhoE = self.J / self.m * self.rho


# number: (label, t, n, m, r, Q [J/mol], V [m^3/mol], A [Pa^-n m^m s^-1], f [Pa], a)
LINEAR_LAWS = {
    15: ("Calcite diffusion creep, Herwegh 2003", 1, 1.1, 3.3, 0.0, 200e3, 0.0, 1.7119e-19, 0.0, 0.0),
    19: ("Anorthite 100 (0.007 wt% H2O), Rybacki & Dresen 2000, 2004", 1, 1.0, 3.0, 0.0, 170e3, 0.0, 5.0119e-23,
         0.0, 0.0),
    25: ("Plagioclase diffusion creep, Rybacki & Dresen 2000", 1, 1.0, 3.0, 0.0, 268e3, 0.0, 1.2589e-19, 0.0, 0.0),
    40: ("Olivine dry diffusion creep, Hirth & Kohlstedt 2003", 1, 1.0, 3.0, 0.0, 375e3, 4.0e-6, 1.5000e-15, 0.0,
         0.0),
    41: ("Olivine wet diffusion creep, Hirth & Kohlstedt 2003", 1, 1.0, 3.0, 1.0, 375e3, 7e-6, 2.5000e-23, 1e9,
         0.0),
}

# number: (label, t, n, m, r, Q, V, A, f, a), as LINEAR_LAWS
GBS_LAWS = {
    40: ("Olivine GBS T < 1250 C, Hirth & Kohlstedt 2003", 1, 3.5, 2.0, 0.0, 400e3, 0.0, 6.5000e-30, 0.0, 0.0),
    41: ("Olivine GBS T > 1250 C, Hirth & Kohlstedt 2003", 1, 3.5, 2.0, 0.0, 600e3, 0.0, 4.7000e-23, 0.0, 0.0),
}

# number: (label, t, q, G (gamma), S [Pa], Q [J/mol], V [m^3/mol], E [Pa^-n s^-1], n)
EXP_LAWS = {
    25: ("Plagioclase Peierls creep, Azuma et al. 2014, regularised as Kameyama et al. 1999", 1, 2.0, 0.2, 9.831e9,
         431.0e3, 0.0, 3.02e-9, 2.0),
    40: ("Olivine Peierls creep, Evans & Goetze 1979, regularised as Kameyama et al. 1999", 1, 2.0, 0.1, 8.5e9,
         5.4e5, 0.0, 5.7e11, 0.0),
}


def correction_factor(t, n):
    """F of ReadDataPowerLaw (FlowLaws.c:577-579): 1 for laws already in
    invariant form, the axial-compression and simple-shear conversions
    otherwise."""
    if t == 0:
        return 1.0
    if t == 1:
        return 1.0 / 6.0 * math.pow(2.0, 1.0 / n) * math.pow(3.0, (n - 1.0) / 2.0 / n)
    if t == 2:
        return 1.0 / 4 * math.pow(2.0, 1.0 / n)
    raise ValueError("experiment type must be 0, 1 or 2")


def power_law(number, scaling, n=1.0, Q=0.0, A=0.0, a=0.0, t=0.0, eta0=0.0, eps0=0.0, tau0=0.0,
              force_act_vol_ast=0, act_vol_dis_ast=0.0, force_melt_weak=0, melt_weak=0.0):
    """ReadDataPowerLaw for phase parameter pwlv = number.  n, Q, A, a, t,
    eta0, eps0, tau0 are the phase's own input values as ReadInputFile left
    them (eta0, eps0, tau0 already divided by their scales; n, Q, A, a, t
    as written), used by the user-defined laws 1-3.  A negative number
    keeps the law but drops the correction factor (F = 1).  Returns the
    scaled t, n, m, r, Q, V, A, f, a and F."""
    k = abs(int(number))
    if k == 1:     # eta = eta0 * exp(-Q / n / R / T) * Eii^(1/n - 1): eta0 back to Pa s, A = eta0^-n
        law = (0.0, n, 0.0, 0.0, Q, 0.0, math.pow(eta0 * scaling.eta, -n), 0.0, 0.0)
    elif k == 2:   # reference stress tau0 at strain rate eps0 (the reference mixes scaled eps0/tau0 here)
        law = (0.0, n, 0.0, 0.0, Q, 0.0, math.pow(math.pow(2.0, n) * eps0 / math.pow(tau0, n), -n), 0.0, 0.0)
    elif k == 3:   # A, n, Q and a as given
        law = (0.0, n, 0.0, 0.0, Q, 0.0, A, 0.0, a)
    elif k in POWER_LAWS:
        law = POWER_LAWS[k][1:]
        if k == 40 and force_act_vol_ast == 1:
            law = law[:5] + (act_vol_dis_ast,) + law[6:]
    else:
        raise ValueError("non-existing power-law flow law number %d" % number)
    t_, n_, m_, r_, Q_, V_, A_, f_, a_ = (float(v) for v in law)
    if force_melt_weak == 1:
        a_ = melt_weak
    S, L = scaling.S, scaling.L
    out = {
        "t": t_, "n": n_, "m": m_, "r": r_,
        "Q": Q_ / scaling.J,
        "V": V_ / L ** 3,
        "A": A_ / (math.pow(S, -n_) * math.pow(S, -r_) * math.pow(L, m_) / scaling.t),
        "f": f_ / S,
        "a": a_,
    }
    out["F"] = 1.0 if number < 0 else correction_factor(t_, n_)
    return out


def _scaled_creep(law, number, scaling):
    t_, n_, m_, r_, Q_, V_, A_, f_, a_ = (float(v) for v in law)
    S, L = scaling.S, scaling.L
    return {"t": t_, "n": n_, "m": m_, "r": r_, "Q": Q_ / scaling.J, "V": V_ / L ** 3,
            "A": A_ / (math.pow(S, -n_) * math.pow(S, -r_) * math.pow(L, m_) / scaling.t),
            "f": f_ / S, "a": a_, "F": 1.0 if number < 0 else correction_factor(t_, n_)}


def linear_law(number, scaling, force_act_vol_ast=0, act_vol_dif_ast=0.0):
    """ReadDataLinear for phase parameter linv = number: the scaled t, n, m,
    r, Q, V, A, f, a and F (a negative number drops the correction factor).
    force_act_vol_ast replaces the activation volume of law 40."""
    k = abs(int(number))
    if k not in LINEAR_LAWS:
        raise ValueError("non-existing linear flow law number %d" % number)
    law = LINEAR_LAWS[k][1:]
    if k == 40 and force_act_vol_ast == 1:
        law = law[:5] + (act_vol_dif_ast,) + law[6:]
    return _scaled_creep(law, number, scaling)


def gbs_law(number, scaling):
    """ReadDataGBS for phase parameter gbsv = number (fields as linear_law)."""
    k = abs(int(number))
    if k not in GBS_LAWS:
        raise ValueError("non-existing GBS flow law number %d" % number)
    return _scaled_creep(GBS_LAWS[k][1:], number, scaling)


def exponential_law(number, scaling):
    """ReadDataExponential for phase parameter expv = number: the scaled
    t, q, G, S, Q, V, E, n (a negative number sets t = 0, which the
    viscosity does not use)."""
    k = abs(int(number))
    if k not in EXP_LAWS:
        raise ValueError("non-existing exponential flow law number %d" % number)
    t_, q_, G_, S_, Q_, V_, E_, n_ = (float(v) for v in EXP_LAWS[k][1:])
    return {"t": 0.0 if number < 0 else t_, "q": q_, "G": G_, "S": S_ / scaling.S, "Q": Q_ / scaling.J,
            "V": V_ / scaling.L ** 3, "E": E_ / (scaling.E * math.pow(scaling.S, -n_)), "n": n_}


def phase_prop_fields(law, pref_pwl=1.0):
    """The power-law fields of mdg_phase_prop for a law from power_law():
    B = pref * F * A^(-1/n) * exp((Q + P V) / (n R T)) * d^(m/n) * f^(-r/n) * exp(-a phi / n)
    with the node's grain size d and melt fraction phi."""
    return {"pwlv": 1, "pref_pwl": pref_pwl, "F_pwl": law["F"], "A_pwl": law["A"], "n_pwl": law["n"],
            "Q_pwl": law["Q"], "V_pwl": law["V"], "m_pwl": law["m"], "r_pwl": law["r"], "f_pwl": law["f"],
            "a_pwl": law["a"]}


def linear_prop_fields(law):
    """The diffusion-creep fields of mdg_phase_prop for a law from linear_law()."""
    out = {"linv": 1}
    out.update({k + "_lin": law[k] for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")})
    return out


def gbs_prop_fields(law):
    """The grain-boundary-sliding fields of mdg_phase_prop for a law from gbs_law()."""
    out = {"gbsv": 1}
    out.update({k + "_gbs": law[k] for k in ("F", "A", "n", "m", "r", "f", "a", "Q", "V")})
    return out


def exponential_prop_fields(law):
    """The Peierls-creep fields of mdg_phase_prop for a law from exponential_law()."""
    return {"expv": 1, "G_exp": law["G"], "S_exp": law["S"], "Q_exp": law["Q"], "E_exp": law["E"],
            "n_exp": law["n"]}


def viscosity_si(law_si, Eii, T, P=0.0, R=8.314, pref=1.0):
    """Invariant-form viscosity [Pa s] of a tabulated law in SI units at
    strain-rate invariant Eii [1/s], temperature T [K], pressure P [Pa]."""
    t, n, m, r, Q, V, A, f, a = law_si
    B = pref * correction_factor(t, n) * math.pow(A, -1.0 / n) * math.exp((Q + P * V) / (n * R * T))
    if r != 0.0:
        B *= math.pow(f, -r / n)
    return B * math.pow(Eii, 1.0 / n - 1.0)
