// Device restatement of ViscosityConcise (RheologyDensity.c:385-1080) for
// constant, power-law, diffusion, grain-boundary-sliding and Peierls creep,
// Maxwell elasticity and Drucker-Prager (plast = 1) phases, shared by the
// grid update and the tangent (Newton) coefficients.
#pragma once
#include "mdg_internal.h"

namespace rheo {

struct rheo_ctl {
    int    step, elastic, average, nph;
    double dt, R, zeroC, min_eta, max_eta, eta_tol, Tiimin, scale_E, scale_eta, T_peierls, d_min;
};

// The creep laws active at one (node, phase) and their coefficients: strain
// rate C T^n of each viscous law at deviatoric stress invariant T.
struct creep_laws {
    bool   elastic, constant, dislocation, peierls, diffusion;
    double eta_el, eta_cst, C_pwl, n_pwl, C_exp, nST, C_lin, n_lin, dm_lin;
};

// LocalIterationViscoElastic (RheologyDensity.c:152-197): Newton iterations on
// eta_ve so that E = el*T/(2 eta_el) + sum of the viscous strain rates,
// T = 2 eta_ve E.  The grain-boundary-sliding term of the reference is
// C_gbs T^n_gbs with C_gbs never set (zero), so it is absent here.  Returns
// false when the reference would exit (20 iterations, res > tol).
__device__ inline bool local_ve(const creep_laws &L, double Eii, double &eta, double &Eii_pwl,
                                double &Eii_exp, double &Eii_lin) {
    const double tol = 1.0e-11;
    for (int it = 0; it < 20; it++) {
        const double Tii = 2.0 * eta * Eii;
        const double e_cst = L.constant ? Tii / 2.0 / L.eta_cst : 0.0;
        const double e_pwl = L.dislocation ? L.C_pwl * pow(Tii, L.n_pwl) : 0.0;
        const double e_exp = L.peierls ? L.C_exp * pow(Tii, L.nST) : 0.0;
        const double e_lin = L.diffusion ? L.C_lin * pow(Tii, L.n_lin) * L.dm_lin : 0.0;
        Eii_pwl = e_pwl;
        Eii_exp = e_exp;
        Eii_lin = e_lin;
        const double r = Eii - (L.elastic ? 1.0 : 0.0) * Tii / (2.0 * L.eta_el) - (e_pwl + e_exp + e_lin + e_cst);
        const double res = fabs(r / Eii);
        if (res < tol / 100) return true;
        if (it == 19 && res > tol) return false;
        double dfdeta = 0.0;
        if (L.elastic) dfdeta += -Eii / L.eta_el;
        if (L.peierls) dfdeta += -e_exp * L.nST / eta;
        if (L.diffusion) dfdeta += -e_lin * L.n_lin / eta;
        if (L.dislocation) dfdeta += -e_pwl * L.n_pwl / eta;
        if (L.constant) dfdeta += -Eii / L.eta_cst;
        eta -= r / dfdeta;
    }
    return true;
}

// Failure codes of viscosity_concise, or-ed into the caller's flag.
constexpr int FAIL_LOCAL_ITER = 1;     // local iterations failed (reference exits, :174-177)
constexpr int FAIL_GRAIN_SIZE = 2;     // d < d_min with m_lin > 0 (reference exits, :506-509)

// ViscosityConcise (RheologyDensity.c:385-1080) for one (node, phase) with
// grain size d and melt fraction phi: returns the dissipative viscosity;
// etaVE, Pcorr, Eii_pl by reference.  EXT = false compiles the diffusion,
// GBS and Peierls laws and the power law's d, f, phi factors out (the
// caller launches it when no phase uses them: the factors are then exactly
// 1 and the laws absent, so the results are the same bits), which keeps
// the register count, and the occupancy, of the constant + power-law case.
template <bool EXT>
__device__ inline double viscosity_concise(const mdg_phase_prop &m, const rheo_ctl &md, double G, double T,
                                           double P, double d, double phi, double Exx, double Ezz, double Exz,
                                           double fric, double C, double dil, double pref_mult, double &etaVE,
                                           double &Pcorr, double &Eii_pl, int &fail) {
    const double dt = (md.step == 0) ? 1e100 : md.dt;
    bool elastic = md.elastic != 0, plastic = (m.plast == 1);
    if (md.step == 0) { elastic = false; plastic = false; }          // :483
    const bool constant = m.cstv != 0, dislocation = m.pwlv != 0;
    const bool peierls = EXT && m.expv != 0 && T < md.T_peierls;      // :474, before the T floor
    const bool diffusion = EXT && m.linv != 0, gbs = EXT && m.gbsv != 0;
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
        if (EXT)
            B_pwl = B_pwl * pow(d, m.m_pwl / m.n_pwl) * pow(m.f_pwl, -m.r_pwl / m.n_pwl) *
                    exp(-m.a_pwl * phi / m.n_pwl);
        C_pwl = pow(2.0 * B_pwl, -m.n_pwl);
    }
    double B_lin = 0.0, C_lin = 0.0;
    if (diffusion) {                                                   // :505-512, d enters below
        if (m.m_lin > 0.0 && d < md.d_min) fail |= FAIL_GRAIN_SIZE;
        B_lin = m.F_lin * pow(m.A_lin, -1.0 / m.n_lin) * exp((m.Q_lin + P * m.V_lin) / md.R / m.n_lin / T) *
                pow(m.f_lin, -m.r_lin / m.n_lin) * exp(-m.a_lin * phi / m.n_lin);
        C_lin = pow(2.0 * B_lin, -m.n_lin);
    }
    double B_gbs = 0.0;
    if (gbs)                                                           // :513-515
        B_gbs = m.F_gbs * pow(m.A_gbs, -1.0 / m.n_gbs) * exp((m.Q_gbs + P * m.V_gbs) / md.R / m.n_gbs / T) *
                pow(d, m.m_gbs / m.n_gbs) * pow(m.f_gbs, -m.r_gbs / m.n_gbs) * exp(-m.a_gbs * phi / m.n_gbs);
    double ST = 0.0, B_exp = 0.0, C_exp = 0.0;
    if (peierls) {                                                     // :516-523
        const double g = m.G_exp;
        ST = m.Q_exp / md.R / T * 2.0 * g * (1 - g);
        const double Arr = exp(-m.Q_exp / md.R / T * pow(1.0 - g, 2.0));
        const double F_exp = pow(pow(2.0, 1.0 - ST - m.n_exp) / pow(sqrt(3.0), ST + m.n_exp + 1.0),
                                 1.0 / (ST + m.n_exp));
        B_exp = F_exp * (pow(g * m.S_exp, ST / (ST + m.n_exp)) / pow(m.E_exp * Arr, 1.0 / (ST + m.n_exp)));
        C_exp = pow(2.0 * B_exp, -(ST + m.n_exp));
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
    if (peierls) eta_up = fmin(eta_up, B_exp * pow(Eii, 1.0 / (ST + m.n_exp) - 1.0));
    if (diffusion) eta_up = fmin(eta_up, B_lin * pow(Eii, 1.0 / m.n_lin - 1.0) * pow(d, m.m_lin / m.n_lin));
    if (gbs) eta_up = fmin(eta_up, B_gbs * pow(Eii, 1.0 / m.n_gbs - 1.0));
    double eta_ve = eta_up, Eii_pwl = 0.0, Eii_exp = 0.0, Eii_lin = 0.0;
    const creep_laws L = {elastic, constant, dislocation, peierls, diffusion, eta_el, eta_cst, C_pwl, m.n_pwl,
                          C_exp, ST + m.n_exp, C_lin, m.n_lin, diffusion ? pow(d, -m.m_lin) : 1.0};
    if (!local_ve(L, Eii, eta_ve, Eii_pwl, Eii_exp, Eii_lin)) fail |= FAIL_LOCAL_ITER;
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
    etaVE = is_pl ? eta_vep : eta_ve;                                  // :962-970
    Eii_pl = is_pl ? gdot / 2.0 : 0.0;
    // dissipative viscosity (:987-1032): each law's viscosity at its own
    // strain-rate share; diffusion creep at the phase's reference grain size
    double inv = 0.0;
    if (peierls) inv += 1.0 / (B_exp * pow(Eii_exp, 1.0 / (ST + m.n_exp) - 1.0));
    if (dislocation) inv += 1.0 / (B_pwl * pow(Eii_pwl, 1.0 / m.n_pwl - 1.0));
    if (diffusion) inv += 1.0 / (B_lin * pow(Eii_lin, 1.0 / m.n_lin - 1.0) * pow(m.gs_ref, m.m_lin / m.n_lin));
    if (constant) inv += 1.0 / eta_cst;
    if (is_pl) inv += 1.0 / (Tii / gdot);
    double eta = 1.0 / inv;
    etaVE = fmax(fmin(etaVE, md.max_eta), md.min_eta);                 // :1058-1073
    if (eta > md.max_eta) eta = md.max_eta;
    if (eta < md.min_eta) eta = md.min_eta;
    return eta;
}

// Whether any phase needs the EXT instantiation of viscosity_concise.
inline bool needs_ext(const mdg_phase_prop *ph, int nph) {
    for (int p = 0; p < nph; p++)
        if (ph[p].linv || ph[p].gbsv || ph[p].expv || ph[p].m_pwl != 0.0 || ph[p].r_pwl != 0.0 ||
            ph[p].a_pwl != 0.0)
            return true;
    return false;
}

// The kernel-side constants of one call.
inline rheo_ctl make_ctl(const mdg_rheo_params *prm, int nph) {
    return {prm->step,     prm->elastic, prm->average, nph,           prm->dt,
            prm->R,        prm->zeroC,   prm->min_eta, prm->max_eta,  prm->eta_tol,
            prm->Tiimin,   prm->scale_E, prm->scale_eta, prm->T_peierls, prm->d_min};
}

// Status of a call from the or-ed failure codes of its nodes.
inline int fail_status(mdg_ctx *c, int fail) {
    if (fail & FAIL_GRAIN_SIZE)
        return mdg_fail(c, MDG_ERR_ARG, "grain size below d_min with grain-size dependent diffusion creep");
    if (fail & FAIL_LOCAL_ITER)
        return mdg_fail(c, MDG_ERR_DIVERGED, "visco-elastic local iterations failed to converge");
    return MDG_OK;
}

}  // namespace rheo
