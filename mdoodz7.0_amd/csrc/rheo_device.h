// Device restatement of ViscosityConcise (RheologyDensity.c:385-1080) for
// constant, power-law, Maxwell-elastic and Drucker-Prager (plast = 1)
// phases, shared by the grid update and the tangent (Newton) coefficients.
#pragma once
#include "mdg_internal.h"

namespace rheo {

struct rheo_ctl {
    int    step, elastic, average, nph;
    double dt, R, zeroC, min_eta, max_eta, eta_tol, Tiimin, scale_E, scale_eta;
};

// LocalIterationViscoElastic (RheologyDensity.c:152-197): Newton iterations on
// eta_ve so that E = el*T/(2 eta_el) + T/(2 eta_cst) + C_pwl T^n, T = 2 eta_ve E.
// Returns false when the reference would exit (20 iterations, res > tol).
__device__ inline bool local_ve(double Eii, double &eta, bool elastic, double eta_el, bool constant,
                         double eta_cst, bool dislocation, double C_pwl, double n_pwl,
                         double &Eii_pwl) {
    const double tol = 1.0e-11;
    for (int it = 0; it < 20; it++) {
        const double Tii = 2.0 * eta * Eii;
        const double e_cst = constant ? Tii / 2.0 / eta_cst : 0.0;
        const double e_pwl = dislocation ? C_pwl * pow(Tii, n_pwl) : 0.0;
        Eii_pwl = e_pwl;
        const double r = Eii - (elastic ? 1.0 : 0.0) * Tii / (2.0 * eta_el) - (e_pwl + e_cst);
        const double res = fabs(r / Eii);
        if (res < tol / 100) return true;
        if (it == 19 && res > tol) return false;
        double dfdeta = 0.0;
        if (elastic) dfdeta += -Eii / eta_el;
        if (dislocation) dfdeta += -e_pwl * n_pwl / eta;
        if (constant) dfdeta += -Eii / eta_cst;
        eta -= r / dfdeta;
    }
    return true;
}

// ViscosityConcise (RheologyDensity.c:385-1080) for one (node, phase):
// returns the dissipative viscosity; etaVE, Pcorr, Eii_pl by reference.
__device__ inline double viscosity_concise(const mdg_phase_prop &m, const rheo_ctl &md, double G, double T,
                                    double P, double Exx, double Ezz, double Exz, double fric,
                                    double C, double dil, double pref_mult, double &etaVE,
                                    double &Pcorr, double &Eii_pl, bool &ok) {
    const double dt = (md.step == 0) ? 1e100 : md.dt;
    bool elastic = md.elastic != 0, plastic = (m.plast == 1);
    if (md.step == 0) { elastic = false; plastic = false; }          // :483
    const bool constant = m.cstv != 0, dislocation = m.pwlv != 0;
    double K = 1.0 / m.beta;
    const double Eyy = -(Exx + Ezz);
    double Eii = sqrt(0.5 * (Exx * Exx + Ezz * Ezz + Eyy * Eyy) + Exz * Exz);   // :457-459
    if (Eii * md.scale_E < 1e-30) Eii = 1e-30 / md.scale_E;
    Pcorr = P;
    if (!elastic) { G = 1e1; dil = 0.0; K = 1e10; }                   // :493
    if (T < md.zeroC) T = md.zeroC;
    double B_pwl = 0.0, C_pwl = 0.0;
    if (dislocation) {                                                 // :501-504
        B_pwl = m.pref_pwl * pref_mult * m.F_pwl * pow(m.A_pwl, -1.0 / m.n_pwl) *
                exp((m.Q_pwl + P * m.V_pwl) / md.R / m.n_pwl / T);
        C_pwl = pow(2.0 * B_pwl, -m.n_pwl);
    }
    const double sin_fric = sin(fric), cos_fric = cos(fric), sin_dil = sin(dil);
    double Tyield = 0.0;
    if (plastic) {                                                     // :530-539
        Tyield = C * cos_fric + P * sin_fric;
        if (m.Slim < Tyield) Tyield = m.Slim;
    }
    const double eta_el = G * dt;
    const double eta_cst = constant ? m.eta0 : 0.0;
    double eta_up = 1.0e100 / md.scale_eta;                            // :585-592
    if (constant) eta_up = fmin(eta_up, eta_cst);
    if (dislocation) eta_up = fmin(eta_up, B_pwl * pow(Eii, 1.0 / m.n_pwl - 1.0));
    if (elastic) eta_up = fmin(eta_up, eta_el);
    double eta_ve = eta_up, Eii_pwl = 0.0;
    if (!local_ve(Eii, eta_ve, elastic, eta_el, constant, eta_cst, dislocation, C_pwl, m.n_pwl,
                  Eii_pwl))
        ok = false;
    double Tii = 2.0 * eta_ve * Eii;
    bool is_pl = false;
    double eta_vep = eta_ve, gdot = 0.0;
    if (plastic) {                                                     // :655-707
        double F_trial = Tii - Tyield;
        if (F_trial > 1e-17) {
            is_pl = true;
            const double F0 = F_trial, tol = md.eta_tol;
            const double Kdt_ss = K * dt * sin_fric * sin_dil;
            double eta_vp = m.eta_vp * pow(Eii, 1.0 / m.n_vp - 1);
            gdot = F_trial / (eta_ve + eta_vp + Kdt_ss);
            double Pc = P, Tiic = Tii;
            for (int it = 0; it < 20; it++) {
                Pc = P + K * dt * (gdot * sin_dil);
                eta_vp = m.eta_vp * pow(fabs(gdot), 1.0 / m.n_vp - 1.0);
                Tyield = Tyield + gdot * eta_vp;    // accumulates, as in the reference
                Tiic = Tii - eta_ve * gdot;
                F_trial = Tiic - Tyield;
                const double res = fabs(F_trial);
                if (res < tol || res / F0 < tol) break;
                gdot -= F_trial / (-eta_ve - eta_vp / m.n_vp - Kdt_ss);
            }
            if (Tiic < 0.0) {                                          // :694-701
                gdot = (Tii - md.Tiimin) / eta_ve;
                Tiic = md.Tiimin;
                Pc = P;
            }
            Pcorr = Pc;
            eta_vep = Tiic / (2.0 * Eii);
            Tii = Tiic;
        }
    }
    etaVE = is_pl ? eta_vep : eta_ve;                                  // :1293-1301
    Eii_pl = is_pl ? gdot / 2.0 : 0.0;
    double inv = 0.0;                                                  // :1024-1030
    if (dislocation) inv += 1.0 / (B_pwl * pow(Eii_pwl, 1.0 / m.n_pwl - 1.0));
    if (constant) inv += 1.0 / eta_cst;
    if (is_pl) inv += 1.0 / (Tii / gdot);
    double eta = 1.0 / inv;
    etaVE = fmax(fmin(etaVE, md.max_eta), md.min_eta);                 // :1058-1073
    if (eta > md.max_eta) eta = md.max_eta;
    if (eta < md.min_eta) eta = md.min_eta;
    return eta;
}

}  // namespace rheo
