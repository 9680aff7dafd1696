/*
 * Oracle: per-node visco-elasto-plastic rheology.  TEST INFRASTRUCTURE ONLY.
 *
 * ViscosityConcise (RheologyDensity.c:385-1080) for constant viscosity,
 * dislocation (power-law) creep with its grain-size, fugacity and melt
 * factors, diffusion (linear) creep, grain-boundary sliding, regularised
 * Peierls creep, Maxwell elasticity and Drucker-Prager plasticity (plast = 1,
 * viscoplastic regularisation eta_vp), with LocalIterationViscoElastic
 * (RheologyDensity.c:152-197), and NonNewtonianViscosityGrid
 * (RheologyDensity.c:1131-1426) for the phase-weighted grid update.
 * Grain-size evolution (gs), Popov plasticity (plast = 2), kinetics,
 * chemical mixing and melting branches are not part of this restatement.
 */
#include <math.h>
#include <stdlib.h>
#include "mdoracle.h"

#define MINV(a, b) ((a) < (b) ? (a) : (b))
#define MAXV(a, b) ((a) > (b) ? (a) : (b))

/* RheologyDensity.c:152-197: Newton iterations on eta_ve so that
 * E = elastic*T/(2 eta_el) + Eii_pwl + Eii_exp + Eii_lin + Eii_gbs + Eii_cst
 * with T = 2 eta_ve E.  The reference's GBS term is C_gbs T^n_gbs with
 * C_gbs left at zero by ViscosityConcise's initialiser (:613-642), so it is
 * zero here. */
typedef struct {
    int    elastic, constant, dislocation, peierls, diffusion;
    double eta_el, eta_cst, C_pwl, n_pwl, C_exp, ST, n_exp, C_lin, n_lin, m_lin, d;
} ve_laws;

static int local_ve(const ve_laws *L, double Eii, double *eta_ve, double *Eii_pwl_out,
                    double *Eii_exp_out, double *Eii_lin_out) {
    const int nitmax = 20;
    const double tol = 1.0e-11;
    double eta = *eta_ve;
    for (int it = 0; it < nitmax; it++) {
        const double Tii = 2.0 * eta * Eii;
        double Eii_cst = 0.0, Eii_pwl = 0.0, Eii_exp = 0.0, Eii_lin = 0.0, Eii_gbs = 0.0;
        if (L->constant) Eii_cst = Tii / 2.0 / L->eta_cst;
        if (L->dislocation) Eii_pwl = L->C_pwl * pow(Tii, L->n_pwl);
        if (L->peierls) Eii_exp = L->C_exp * pow(Tii, L->ST + L->n_exp);
        if (L->diffusion) Eii_lin = L->C_lin * pow(Tii, L->n_lin) * pow(L->d, -L->m_lin);
        const double Eii_vis = Eii_pwl + Eii_exp + Eii_lin + Eii_gbs + Eii_cst;
        const double r = Eii - L->elastic * Tii / (2.0 * L->eta_el) - Eii_vis;
        const double res = fabs(r / Eii);
        *Eii_pwl_out = Eii_pwl;
        *Eii_exp_out = Eii_exp;
        *Eii_lin_out = Eii_lin;
        if (res < tol / 100) break;
        if (it == nitmax - 1 && res > tol) return -1;   /* reference exits */
        double dfdeta = 0.0;
        if (L->elastic) dfdeta += -Eii / L->eta_el;
        if (L->peierls) dfdeta += -Eii_exp * (L->ST + L->n_exp) / eta;
        if (L->diffusion) dfdeta += -Eii_lin * L->n_lin / eta;
        if (L->dislocation) dfdeta += -Eii_pwl * L->n_pwl / eta;
        if (L->constant) dfdeta += -Eii / L->eta_cst;
        eta -= r / dfdeta;
    }
    *eta_ve = eta;
    return 0;
}

/* One phase at one node: returns the effective (VE or VEP) viscosity in
 * *etaVE, the dissipative viscosity as return value, and Pcorr. */
double orc_viscosity_concise(const orc_phase *m, const orc_rheo_model *md, double G, double T,
                             double P, double d, double phi, double Exx, double Ezz, double Exz,
                             double fric, double C, double dil, double pref_mult, double *etaVE,
                             double *Pcorr, double *Eii_pl, int *status) {
    double dt = (md->step == 0) ? 1e100 : md->dt;
    int elastic = md->elastic, plastic = (m->plast == 1);
    const int constant = m->cstv != 0, dislocation = m->pwlv != 0;
    const int peierls = m->expv != 0 && T < md->T_peierls;   /* :474, before the T floor */
    const int diffusion = m->linv != 0, gbs = m->gbsv != 0;
    if (md->step == 0) { elastic = 0; plastic = 0; }
    double K = 1.0 / m->beta_for_K;
    const double Eyy = -(Exx + Ezz);
    double Eii = sqrt(0.5 * (Exx * Exx + Ezz * Ezz + Eyy * Eyy) + Exz * Exz);
    if (Eii * md->scale_E < 1e-30) Eii = 1e-30 / md->scale_E;
    *Pcorr = P;
    if (!elastic) { G = 1e1; dil = 0.0; K = 1e10; }
    if (T < md->zeroC) T = md->zeroC;
    /* precomputations, :501-523 */
    double B_pwl = 0.0, C_pwl = 0.0, B_lin = 0.0, C_lin = 0.0, B_gbs = 0.0;
    double ST = 0.0, B_exp = 0.0, C_exp = 0.0;
    if (dislocation) {
        B_pwl = m->pref_pwl * pref_mult * m->F_pwl * pow(m->A_pwl, -1.0 / m->n_pwl) *
                exp((m->Q_pwl + P * m->V_pwl) / md->R / m->n_pwl / T) * pow(d, m->m_pwl / m->n_pwl) *
                pow(m->f_pwl, -m->r_pwl / m->n_pwl) * exp(-m->a_pwl * phi / m->n_pwl);
        C_pwl = pow(2.0 * B_pwl, -m->n_pwl);
    }
    if (diffusion) {
        if (m->m_lin > 0.0 && d < md->d_min) *status = -2;   /* reference exits */
        B_lin = m->F_lin * pow(m->A_lin, -1.0 / m->n_lin) * exp((m->Q_lin + P * m->V_lin) / md->R / m->n_lin / T) *
                pow(m->f_lin, -m->r_lin / m->n_lin) * exp(-m->a_lin * phi / m->n_lin);
        C_lin = pow(2.0 * B_lin, -m->n_lin);
    }
    if (gbs)
        B_gbs = m->F_gbs * pow(m->A_gbs, -1.0 / m->n_gbs) * exp((m->Q_gbs + P * m->V_gbs) / md->R / m->n_gbs / T) *
                pow(d, m->m_gbs / m->n_gbs) * pow(m->f_gbs, -m->r_gbs / m->n_gbs) * exp(-m->a_gbs * phi / m->n_gbs);
    if (peierls) {
        const double gamma = m->G_exp, n_exp = m->n_exp;
        ST = m->Q_exp / md->R / T * 2.0 * gamma * (1 - gamma);
        const double Arr_exp = exp(-m->Q_exp / md->R / T * pow(1.0 - gamma, 2.0));
        const double F_exp = pow(pow(2.0, 1.0 - ST - n_exp) / pow(sqrt(3.0), ST + n_exp + 1.0), 1.0 / (ST + n_exp));
        B_exp = F_exp * (pow(gamma * m->S_exp, ST / (ST + n_exp)) / pow(m->E_exp * Arr_exp, 1.0 / (ST + n_exp)));
        C_exp = pow(2.0 * B_exp, -(ST + n_exp));
    }
    const double cos_fric = cos(fric), sin_fric = sin(fric), sin_dil = sin(dil);
    double Tyield = 0.0;
    if (plastic) {
        Tyield = C * cos_fric + P * sin_fric;
        if (m->Slim < Tyield) { C = m->Slim; Tyield = m->Slim; }
    }
    /* isolated viscosities, :559-580 */
    const double eta_el = G * dt;
    double eta_cst = 0.0, eta_pwl = 0.0, eta_lin = 0.0, eta_gbs = 0.0, eta_exp = 0.0;
    if (constant) eta_cst = m->eta0;
    if (dislocation) eta_pwl = B_pwl * pow(Eii, 1.0 / m->n_pwl - 1.0);
    if (diffusion) eta_lin = B_lin * pow(Eii, 1.0 / m->n_lin - 1.0) * pow(d, m->m_lin / m->n_lin);
    if (gbs) eta_gbs = B_gbs * pow(Eii, 1.0 / m->n_gbs - 1.0);
    if (peierls) eta_exp = B_exp * pow(Eii, 1.0 / (ST + m->n_exp) - 1.0);
    /* upper bound of the visco-elastic viscosity, :585-592 */
    double eta_up = 1.0e100 / md->scale_eta;
    if (constant) eta_up = MINV(eta_up, eta_cst);
    if (dislocation) eta_up = MINV(eta_up, eta_pwl);
    if (elastic) eta_up = MINV(eta_up, eta_el);
    if (peierls) eta_up = MINV(eta_up, eta_exp);
    if (diffusion) eta_up = MINV(eta_up, eta_lin);
    if (gbs) eta_up = MINV(eta_up, eta_gbs);
    double eta_ve = eta_up, Eii_pwl = 0.0, Eii_exp = 0.0, Eii_lin = 0.0;
    const ve_laws L = {elastic, constant, dislocation, peierls, diffusion, eta_el, eta_cst, C_pwl, m->n_pwl,
                       C_exp, ST, m->n_exp, C_lin, m->n_lin, m->m_lin, d};
    if (local_ve(&L, Eii, &eta_ve, &Eii_pwl, &Eii_exp, &Eii_lin) && *status == 0)
        *status = -1;
    double Tii = 2.0 * eta_ve * Eii;
    int is_pl = 0;
    double eta_vep = eta_ve, gdot = 0.0, eta_vp = m->eta_vp;
    if (plastic) {
        /* RheologyDensity.c:655-707 */
        double F_trial = Tii - Tyield;
        if (F_trial > 1e-17) {
            is_pl = 1;
            const double F_trial0 = F_trial, tol = md->eta_tol;
            eta_vp = m->eta_vp * pow(Eii, 1.0 / m->n_vp - 1);
            gdot = F_trial / (eta_ve + eta_vp + K * dt * sin_fric * sin_dil);
            double Pc = P, Tiic = Tii;
            for (int it = 0; it < 20; it++) {
                const double A_P = -sin_dil;
                const double divp = -gdot * A_P;
                Pc = P + K * dt * divp;
                eta_vp = m->eta_vp * pow(fabs(gdot), 1.0 / m->n_vp - 1.0);
                Tyield = Tyield + gdot * eta_vp;
                Tiic = Tii - eta_ve * gdot;
                F_trial = Tiic - Tyield;
                const double res = fabs(F_trial);
                if (res < tol || res / F_trial0 < tol) break;
                const double dFdgdot = -eta_ve - eta_vp / m->n_vp - K * dt * sin_fric * sin_dil;
                gdot -= F_trial / dFdgdot;
            }
            if (Tiic < 0.0) {   /* tension: von Mises minimum stress Tiimin */
                gdot = (Tii - md->Tiimin) / eta_ve;
                Tiic = md->Tiimin;
                Pc = P;
            }
            *Pcorr = Pc;
            eta_vep = Tiic / (2.0 * Eii);
            Tii = Tiic;
        }
    }
    *etaVE = is_pl ? eta_vep : eta_ve;
    *Eii_pl = is_pl ? gdot / 2.0 : 0.0;
    /* dissipative viscosity (post-processing branch, RheologyDensity.c:987-1032):
     * each law at its own strain-rate share; diffusion creep at the grain
     * size d1 = gs_ref the reference sets for phases without grain-size
     * evolution (:486) */
    if (dislocation) eta_pwl = B_pwl * pow(Eii_pwl, 1.0 / m->n_pwl - 1.0);
    if (diffusion) eta_lin = B_lin * pow(Eii_lin, 1.0 / m->n_lin - 1.0) * pow(m->gs_ref, m->m_lin / m->n_lin);
    if (peierls) eta_exp = B_exp * pow(Eii_exp, 1.0 / (ST + m->n_exp) - 1.0);
    double inv = 0.0;
    if (peierls) inv += 1.0 / eta_exp;
    if (dislocation) inv += 1.0 / eta_pwl;
    if (diffusion) inv += 1.0 / eta_lin;
    if (constant) inv += 1.0 / eta_cst;
    if (is_pl) inv += 1.0 / (Tii / gdot);
    double eta = 1.0 / inv;
    if (*etaVE > md->max_eta) *etaVE = md->max_eta;
    if (*etaVE < md->min_eta) *etaVE = md->min_eta;
    if (eta > md->max_eta) eta = md->max_eta;
    if (eta < md->min_eta) eta = md->min_eta;
    return eta;
}

/* NonNewtonianViscosityGrid (RheologyDensity.c:1131-1426) on one node set:
 * n nodes, nph phases with fractions perc[p*n + i].  Centres skip tags 30
 * and 31 (:1188), vertices tag 30 (:1333). */
int orc_rheology_nodes(const orc_rheo_model *md, const orc_phase *phases, int nph, int n,
                       const double *perc, const double *G, const double *T, const double *P,
                       const double *exx, const double *ezz, const double *exz,
                       const double *sxx0, const double *szz0, const double *sxz0,
                       const double *fric, const double *C, const double *dil,
                       const double *pref, const double *d, const double *phi,
                       const signed char *tag, int centroid, int average,
                       double *eta_out, double *eta_phys, double *sxx, double *szz, double *sxz,
                       double *p_corr, double *eII_pl) {
    const double el = md->elastic ? 1.0 : 0.0;
    int status = 0;
    for (int i = 0; i < n; i++) {
        if (tag && (tag[i] == 30 || (centroid && tag[i] == 31))) {
            /* air cells / vertices: every output left at zero */
            eta_out[i] = 0.0;
            if (eta_phys) eta_phys[i] = 0.0;
            if (sxx) sxx[i] = 0.0;
            if (szz) szz[i] = 0.0;
            if (sxz) sxz[i] = 0.0;
            if (p_corr) p_corr[i] = 0.0;
            if (eII_pl) eII_pl[i] = 0.0;
            continue;
        }
        const double eta_e = md->elastic ? md->dt * G[i] : 1.0;
        const double Exx = exx[i] + el * sxx0[i] / eta_e / 2.0;
        const double Ezz = ezz[i] + el * szz0[i] / eta_e / 2.0;
        const double Exz = exz[i] + el * sxz0[i] / eta_e / 2.0;
        double e = 0.0, ep = 0.0, pc = 0.0, epl = 0.0;
        for (int p = 0; p < nph; p++) {
            const double f = perc[(size_t)p * n + i];
            if (!(fabs(f) > 1e-13)) continue;
            double etaVE, Pcorr, Eii_pl;
            const double eta = orc_viscosity_concise(&phases[p], md, G[i], T[i], P[i], d ? d[i] : 0.0,
                                                     phi ? phi[i] : 0.0, Exx, Ezz, Exz,
                                                     fric ? fric[i] : phases[p].fric,
                                                     C ? C[i] : phases[p].C,
                                                     dil ? dil[i] : phases[p].dil,
                                                     pref ? pref[i] : 1.0, &etaVE, &Pcorr, &Eii_pl,
                                                     &status);
            if (average == 0) { e += f * etaVE; ep += f * eta; }
            if (average == 1) { e += f / etaVE; ep += f / eta; }
            if (average == 2) { e += f * log(etaVE); ep += f * log(eta); }
            pc += f * Pcorr;
            epl += f * Eii_pl;
        }
        if (average == 1) { e = 1.0 / e; ep = 1.0 / ep; }
        if (average == 2) { e = exp(e); ep = exp(ep); }
        eta_out[i] = e;
        if (eta_phys) eta_phys[i] = ep;
        if (sxx) sxx[i] = 2.0 * e * Exx;
        if (szz) szz[i] = 2.0 * e * Ezz;
        if (sxz) sxz[i] = 2.0 * e * Exz;
        if (p_corr) p_corr[i] = pc;
        if (eII_pl) eII_pl[i] = epl;
    }
    return status;
}

/* Softening (RheologyDensity.c:1432-1535) of one phase at accumulated plastic
 * strain e: returns fric, dil, C through the pointers. */
static void soften(const orc_phase_soft *m, double e, int style, double *fric, double *dil, double *C) {
    *fric = m->phi; *dil = m->psi; *C = m->C;
    const double dstrain = m->pls_end - m->pls_start;
    if (style == 1) {
        const double mu = 0.5 * (m->pls_end + m->pls_start);
        if (m->phi_soft == 1) {
            const double d = m->phi - m->phi_end;
            const double f = m->phi - d / 2.0 * erfc(-(e - mu) / dstrain) - m->phi_end;
            const double f0 = m->phi - d / 2.0 * erfc(-(0.0 - mu) / dstrain) - m->phi_end;
            *fric = f * d / f0 + m->phi_end;
        }
        if (m->psi_soft == 1) {
            const double d = m->psi - m->psi_end;
            const double f = m->psi - d / 2.0 * erfc(-(e - mu) / dstrain) - m->psi_end;
            const double f0 = m->psi - d / 2.0 * erfc(-(0.0 - mu) / dstrain) - m->psi_end;
            *dil = f * d / f0 + m->psi_end;
        }
        if (m->coh_soft == 1) {
            const double d = m->C - m->C_end;
            const double f = m->C - d / 2.0 * erfc(-(e - mu) / dstrain) - m->C_end;
            const double f0 = m->C - d / 2.0 * erfc(-(0.0 - mu) / dstrain) - m->C_end;
            *C = f * d / f0 + m->C_end;
        }
    } else if (style == 2) {
        *C = fmax(m->C + m->H_c * e, m->C_end);
    } else {
        if (e < m->pls_start) {
            if (m->phi_soft == 1) *fric = m->phi;
            if (m->psi_soft == 1) *dil = m->psi;
            if (m->coh_soft == 1) *C = m->C;
        }
        if (e >= m->pls_end) {
            if (m->phi_soft == 1) *fric = m->phi_end;
            if (m->psi_soft == 1) *dil = m->psi_end;
            if (m->coh_soft == 1) *C = m->C_end;
        }
        if (e >= m->pls_start && e < m->pls_end) {
            /* the reference scales the slope by e itself, not e - pls_start */
            if (m->phi_soft == 1) *fric = m->phi + (m->phi_end - m->phi) / dstrain * e;
            if (m->psi_soft == 1) *dil = m->psi + (m->psi_end - m->psi) / dstrain * e;
            if (m->coh_soft == 1) *C = m->C + (m->C_end - m->C) / dstrain * e;
        }
    }
}

/* ShearModCompExpGrid (RheologyDensity.c:1634-1730, harmonic G and beta)
 * and CohesionFrictionDilationGrid (:1539-1616, arithmetic phase average of
 * the softened fric, dil, C) on one node set.  strain_pl may be NULL (no
 * softening outputs). */
void orc_material_grid(const orc_phase_soft *ph, int nph, int n, const double *perc,
                       const double *strain_pl, const signed char *tag, int centroid, int style,
                       double *mu, double *bet, double *C, double *fric, double *dil) {
    for (int i = 0; i < n; i++) {
        mu[i] = bet[i] = 0.0;
        if (strain_pl) C[i] = fric[i] = dil[i] = 0.0;
        if (tag && (tag[i] == 30 || (centroid && tag[i] == 31))) continue;
        double a = 0.0, b = 0.0, cc = 0.0, ff = 0.0, dd = 0.0;
        for (int p = 0; p < nph; p++) {
            const double w = perc[(size_t)p * n + i];
            a += w * 1.0 / ph[p].G;
            b += w * 1.0 / ph[p].bet;
            if (strain_pl) {
                double f, d, c;
                soften(&ph[p], strain_pl[i], style, &f, &d, &c);
                ff += w * f; dd += w * d; cc += w * c;
            }
        }
        mu[i] = 1.0 / a;
        bet[i] = 1.0 / b;
        if (strain_pl) { C[i] = cc; fric[i] = ff; dil[i] = dd; }
    }
}

/* Newton tangent coefficients (RheologicalOperators with Jacobian = 1,
 * StokesRoutines.c:250-375, isotropic and incompressible) with the
 * derivatives of the phase-averaged effective viscosity by central finite
 * differences (DerivativesOnTheFly_n/_s, PhaseRheologyLoop_v1,
 * FD_Jacobian.c:48-323): perturbation tol*|E| + tol/1e10 with tol = 1e-5
 * on centres and 1e-7 on vertices; the shear derivative is taken w.r.t.
 * gxz = 2 Exz (denominator 4 pert).  D has 8 rows on centres (D11..D14,
 * D21..D24) and 4 on vertices (D31..D34), each of n nodes. */
int orc_tangent_nodes(const orc_rheo_model *md, const orc_phase *phases, int nph, int n,
                      const double *perc, const double *G, const double *T, const double *P,
                      const double *exx, const double *ezz, const double *exz,
                      const double *sxx0, const double *szz0, const double *sxz0,
                      const double *fric, const double *C, const double *dil,
                      const double *pref, const double *d, const double *phi,
                      const signed char *tag, int centroid, int average, double *D) {
    const double el = md->elastic ? 1.0 : 0.0;
    const double tol = centroid ? 1e-5 : 1e-7;
    const int nrow = centroid ? 8 : 4;
    int status = 0;
    for (int i = 0; i < n; i++) {
        for (int r = 0; r < nrow; r++) D[(size_t)r * n + i] = 0.0;
        if (tag && (tag[i] == 30 || (centroid && tag[i] == 31))) continue;
        const double eta_e = md->elastic ? md->dt * G[i] : 1.0;
        /* EffectiveStrainRate (RheologyDensity.c:116-146), isotropic */
        const double Exx = exx[i] + el * sxx0[i] / eta_e / 2.0;
        const double Ezz = ezz[i] + el * szz0[i] / eta_e / 2.0;
        const double Exz = exz[i] + el * sxz0[i] / eta_e / 2.0;
        const double px = tol * fabs(Exx) + tol / 1e10, pz = tol * fabs(Ezz) + tol / 1e10;
        const double pg = tol * fabs(Exz) + tol / 1e10, pp = tol * fabs(P[i]) + tol / 1e10;
        /* perturbed states: (dExx, dEzz, dExz, dP, sign, denominator) */
        const double st[8][6] = {{px, 0, 0, 0, 1, 2 * px}, {-px, 0, 0, 0, -1, 2 * px},
                                 {0, pz, 0, 0, 1, 2 * pz}, {0, -pz, 0, 0, -1, 2 * pz},
                                 {0, 0, pg, 0, 1, 4 * pg}, {0, 0, -pg, 0, -1, 4 * pg},
                                 {0, 0, 0, pp, 1, 2 * pp}, {0, 0, 0, -pp, -1, 2 * pp}};
        double der[4] = {0, 0, 0, 0}, e = 0.0;
        /* phase-averaged effective viscosity and per-phase values at the base state */
        double eta_ph[64];
        for (int p = 0; p < nph; p++) {
            const double f = perc[(size_t)p * n + i];
            eta_ph[p] = 0.0;
            if (!(fabs(f) > 1e-13)) continue;
            double etaVE, Pc, Epl;
            orc_viscosity_concise(&phases[p], md, G[i], T[i], P[i], d ? d[i] : 0.0, phi ? phi[i] : 0.0,
                                  Exx, Ezz, Exz, fric ? fric[i] : phases[p].fric, C ? C[i] : phases[p].C,
                                  dil ? dil[i] : phases[p].dil, pref ? pref[i] : 1.0, &etaVE,
                                  &Pc, &Epl, &status);
            eta_ph[p] = etaVE;
            if (average == 0) e += f * etaVE;
            if (average == 1) e += f / etaVE;
            if (average == 2) e += f * log(etaVE);
        }
        if (average == 1) e = 1.0 / e;
        if (average == 2) e = exp(e);
        /* Central differences per phase.  The reference accumulates the
         * scaled +h and -h terms of all phases into one sum
         * (PhaseRheologyLoop_v1, FD_Jacobian.c:66-78), which leaves round-off
         * of size ulp(eta)/h in cells holding two phases (O(1) with the
         * h = 1e-15 floor of a zero strain-rate component); here each phase's
         * +h and -h values are differenced first: the same central difference
         * without that cancellation. */
        for (int q = 0; q < 4; q++) {
            for (int p = 0; p < nph; p++) {
                const double f = perc[(size_t)p * n + i];
                if (!(fabs(f) > 1e-13)) continue;
                double eta_pm[2];
                for (int sgn = 0; sgn < 2; sgn++) {
                    const int s = 2 * q + sgn;
                    double Pc, Epl;
                    orc_viscosity_concise(&phases[p], md, G[i], T[i], P[i] + st[s][3], d ? d[i] : 0.0,
                                          phi ? phi[i] : 0.0, Exx + st[s][0],
                                          Ezz + st[s][1], Exz + st[s][2],
                                          fric ? fric[i] : phases[p].fric, C ? C[i] : phases[p].C,
                                          dil ? dil[i] : phases[p].dil, pref ? pref[i] : 1.0,
                                          &eta_pm[sgn], &Pc, &Epl, &status);
                }
                double v = f * (eta_pm[0] - eta_pm[1]) / st[2 * q][5];
                if (average == 1) v /= eta_ph[p] * eta_ph[p];
                if (average == 2) v /= eta_ph[p];
                der[q] += v;
            }
        }
        for (int q = 0; q < 4; q++) {
            if (average == 1) der[q] *= e * e;
            if (average == 2) der[q] *= e;
        }
        if (centroid) {
            D[0 * (size_t)n + i] = 2.0 * e + 2.0 * der[0] * Exx;
            D[1 * (size_t)n + i] = 2.0 * der[1] * Exx;
            D[2 * (size_t)n + i] = 2.0 * der[2] * Exx;
            D[3 * (size_t)n + i] = 2.0 * der[3] * Exx;
            D[4 * (size_t)n + i] = 2.0 * der[0] * Ezz;
            D[5 * (size_t)n + i] = 2.0 * e + 2.0 * der[1] * Ezz;
            D[6 * (size_t)n + i] = 2.0 * der[2] * Ezz;
            D[7 * (size_t)n + i] = 2.0 * der[3] * Ezz;
        } else {
            /* Dxz = Gxz/2 = Exz for the isotropic case */
            D[0 * (size_t)n + i] = 2.0 * der[0] * Exz;
            D[1 * (size_t)n + i] = 2.0 * der[1] * Exz;
            D[2 * (size_t)n + i] = e + 2.0 * der[2] * Exz;
            D[3 * (size_t)n + i] = 2.0 * der[3] * Exz;
        }
    }
    return status;
}
