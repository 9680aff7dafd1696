/*
 * Oracle: staggered-grid Stokes residual (Assemble=0) and Picard operator
 * assembly (Assemble=1).  TEST INFRASTRUCTURE ONLY.
 *
 * Supported subset: no free surface (tag 30), no stress BC (tag 2), no
 * periodic tags (-2/-12), no density stabilisation.  These are the
 * boundary conditions of every workload the tests and bench.py use.
 */
#include <math.h>
#include <stdlib.h>
#include "mdoracle.h"

static inline int is_free_u(signed char t) { return t == -1 || t == 2 || t == -2; }
/* AddCoeff3 (StokesAssemblyDecoupled.c:46-56): Dirichlet-like neighbours
 * are folded into the right-hand side instead of the matrix. */
static inline int folds_to_rhs(signed char t) { return t == 0 || t == 31 || t == 11 || t == 13; }

void orc_residual(const orc_grid *g, const int *eqn_u, const int *eqn_v, const int *eqn_p,
                  int neq_mom, const double *sxxd, const double *szzd, const double *sxz,
                  const double *p_corr, const double *div, const double *roger_x,
                  const double *roger_z, const double *rhs_p, const double *d, double *F,
                  double norms[3]) {
    const int nx = g->Nx, nz = g->Nz, nzvx = nz + 1, nxvz = nx + 1, ncx = nx - 1, ncz = nz - 1;
    const double dx = g->dx, dz = g->dz, celvol = dx * dz;
    double resx = 0.0, resz = 0.0, resp = 0.0;
    int ndofx = 0, ndofz = 0, ndofp = 0;
    /* X momentum, StokesAssemblyDecoupled.c:184-211 (regular rows). */
    for (int l = 1; l < nzvx - 1; l++)
        for (int k = 0; k < nx; k++) {
            const int c1 = k + l * nx;
            if (!is_free_u(g->BCu_t[c1])) continue;
            const int c2 = k - 1 + (l - 1) * ncx;
            const int iPrW = c2, iPrE = c2 + 1, ixyN = c1, ixyS = c1 - nx;
            double f = (sxxd[iPrE] - sxxd[iPrW]) / dx;
            f -= (p_corr[iPrE] - p_corr[iPrW]) / dx;
            f += (sxz[ixyN] - sxz[ixyS]) / dz;
            f += roger_x[c1];
            F[eqn_u[c1]] = -celvol * f;
        }
    /* Z momentum, StokesAssemblyDecoupled.c:688-724. */
    for (int l = 0; l < nz; l++)
        for (int k = 1; k < nxvz - 1; k++) {
            const int c3 = k + l * nxvz;
            if (!(g->BCv_t[c3] == -1 || g->BCv_t[c3] == 2)) continue;
            const int c1 = k + l * nx, c2 = k - 1 + (l - 1) * ncx;
            const int iPrS = c2, iPrN = c2 + ncx, ixyE = c1, ixyW = c1 - 1;
            double f = (szzd[iPrN] - szzd[iPrS]) / dz;
            f -= (p_corr[iPrN] - p_corr[iPrS]) / dz;
            f += (sxz[ixyE] - sxz[ixyW]) / dx;
            f += roger_z[c3];
            F[eqn_v[c3]] = -celvol * f;
        }
    /* Continuity, StokesAssemblyDecoupled.c:1128-1144 (incompressible). */
    for (int c = 0; c < ncx * ncz; c++) {
        if (g->BCp_t[c] != -1) continue;
        F[eqn_p[c]] = celvol * (div[c] - (rhs_p ? rhs_p[c] : 0.0));
    }
    /* Norms, StokesRoutines.c:890-932: momentum rows scaled by d first and
     * summed away from the two rows nearest each wall; continuity summed
     * before scaling over every free cell. */
    for (int cc = 0; cc < nzvx * nx; cc++) {
        const signed char t = g->BCu_t[cc];
        if (t == 0 || t == 30 || t == 11 || t == 13 || t == -12) continue;
        const int e = eqn_u[cc], l = cc / nx;
        if (d) F[e] *= d[e];
        if (l > 2 && l < nzvx - 2) { ndofx++; resx += F[e] * F[e]; }
    }
    for (int cc = 0; cc < nz * nxvz; cc++) {
        const signed char t = g->BCv_t[cc];
        if (t == 0 || t == 30 || t == 11 || t == 13 || t == -12) continue;
        const int e = eqn_v[cc], k = cc % nxvz;
        if (d) F[e] *= d[e];
        if (k > 2 && k < nxvz - 2) { ndofz++; resz += F[e] * F[e]; }
    }
    for (int cc = 0; cc < ncz * ncx; cc++) {
        const signed char t = g->BCp_t[cc];
        if (t == 0 || t == 30 || t == 31) continue;
        const int e = eqn_p[cc];
        ndofp++;
        resp += F[e] * F[e];
        if (d) F[e] *= d[e];
    }
    (void)neq_mom;
    norms[0] = sqrt(resx / ndofx);
    norms[1] = sqrt(resz / ndofz);
    norms[2] = sqrt(resp / ndofp);
}

/* Picard stencils, StokesAssemblyDecoupled.c:62-212 (x), 562-725 (z),
 * 1090-1145 (continuity), for SxxBC = SzzBC = 0 and out_of_plane = 0.
 * The sympy-generated coefficients reduce to the forms below; with
 * comp = 0 they are the familiar 2*eta*d/dx, eta*(du/dz + dv/dx) terms. */
void orc_assemble_picard(const orc_grid *g, const int *eqn_u, const int *eqn_v,
                         const int *eqn_p, int neq_mom, int neq_cont, const double *D11,
                         const double *D22, const double *D33,
                         int *Ap, int *Aj, double *Ax, int *Bp, int *Bj, double *Bx,
                         int *Cp, int *Cj, double *Cx) {
    const int nx = g->Nx, nz = g->Nz, nzvx = nz + 1, nxvz = nx + 1, ncx = nx - 1, ncz = nz - 1;
    const double dx = g->dx, dz = g->dz, celvol = dx * dz, comp = 0.0;
    const signed char *Ut = g->BCu_t, *Vt = g->BCv_t, *Pt = g->BCp_t, *Gt = g->BCg_t;
    int na = 0, nb = 0;
    int row = 0;
#define PUTA(col_t, col_e, coef)                                                         \
    do { if (!folds_to_rhs(col_t)) { Aj[na] = (col_e); Ax[na] = (coef) * celvol; na++; } } while (0)
#define PUTB(col_t, col_e, coef)                                                         \
    do { if (!folds_to_rhs(col_t)) { Bj[nb] = (col_e) - neq_mom; Bx[nb] = (coef) * celvol; nb++; } } while (0)
    /* rows in equation order: Vx rows then Vz rows (StokesRoutines.c:1162-1220) */
    for (int l = 0; l < nzvx; l++)
        for (int k = 0; k < nx; k++) {
            const int c1 = k + l * nx;
            if (!is_free_u(Ut[c1])) continue;
            Ap[row] = na; Bp[row] = nb; row++;
            if (!(l > 0 && l < nzvx - 1)) continue;
            const int c2 = k - 1 + (l - 1) * ncx, c3 = k + l * nxvz;
            const int iVxS = c1 - nx, iVxW = c1 - 1, iVxE = c1 + 1, iVxN = c1 + nx;
            const int iVzSW = c3 - nxvz, iVzSE = c3 - nxvz + 1, iVzNW = c3, iVzNE = c3 + 1;
            const int iPrW = c2, iPrE = c2 + 1, ixyN = c1, ixyS = c1 - nx;
            const double D11E = D11[iPrE], D11W = D11[iPrW], D33N = D33[ixyN], D33S = D33[ixyS];
            const double inW = (Pt[iPrW] == -1), inE = (Pt[iPrE] == -1);
            const double inS = (Gt[ixyS] != 30), inN = (Gt[ixyN] != 30);
            double RegS = 1, RegN = 1, DirS = 0, DirN = 0;
            if (Ut[iVxS] == 13) RegS = 0;
            if (Ut[iVxN] == 13) RegN = 0;
            if (Ut[iVxS] == 11) { DirS = 1; RegS = 0; }
            if (Ut[iVxN] == 11) { DirN = 1; RegN = 0; }
            const double Q = D33N * inN * (2 * DirN + RegN) + D33S * inS * (2 * DirS + RegS);
            const double uW = D11W * inW * (comp * inW - 3) / (3.0 * dx * dx);
            const double uE = D11E * inE * (comp * inE - 3) / (3.0 * dx * dx);
            const double uS = -D33S * RegS * inS / (dz * dz);
            const double uN = -D33N * RegN * inN / (dz * dz);
            const double uC = Q / (dz * dz) - (D11E * inE * (comp * inE - 3) +
                                               D11W * inW * (comp * inW - 3)) / (3.0 * dx * dx);
            const double vSW = (D11W * comp * inW * inW - 3 * D33S * inS) / (3.0 * dx * dz);
            const double vSE = -(D11E * comp * inE * inE - 3 * D33S * inS) / (3.0 * dx * dz);
            const double vNW = -(D11W * comp * inW * inW - 3 * D33N * inN) / (3.0 * dx * dz);
            const double vNE = (D11E * comp * inE * inE - 3 * D33N * inN) / (3.0 * dx * dz);
            if (Ut[iVxS] != 30) PUTA(Ut[iVxS], eqn_u[iVxS], uS);
            if (Ut[iVxW] != 30 && k > 0) PUTA(Ut[iVxW], eqn_u[iVxW], uW);
            if (Ut[c1] != 30) PUTA(Ut[c1], eqn_u[c1], uC);
            if (Ut[iVxE] != 30 && k < nx - 1) PUTA(Ut[iVxE], eqn_u[iVxE], uE);
            if (Ut[iVxN] != 30) PUTA(Ut[iVxN], eqn_u[iVxN], uN);
            if (Vt[iVzSW] != 30 && k > 0) PUTA(Vt[iVzSW], eqn_v[iVzSW], vSW);
            if (Vt[iVzSE] != 30 && k < nx - 1) PUTA(Vt[iVzSE], eqn_v[iVzSE], vSE);
            if (Vt[iVzNW] != 30 && k > 0) PUTA(Vt[iVzNW], eqn_v[iVzNW], vNW);
            if (Vt[iVzNE] != 30 && k < nx - 1) PUTA(Vt[iVzNE], eqn_v[iVzNE], vNE);
            if (Pt[iPrW] != 30 && k > 0) PUTB(Pt[iPrW], eqn_p[iPrW], -1.0 / dx);
            if (Pt[iPrE] != 30 && k < nx - 1) PUTB(Pt[iPrE], eqn_p[iPrE], 1.0 / dx);
        }
    for (int l = 0; l < nz; l++)
        for (int k = 0; k < nxvz; k++) {
            const int c3 = k + l * nxvz;
            if (!(Vt[c3] == -1 || Vt[c3] == 2)) continue;
            Ap[row] = na; Bp[row] = nb; row++;
            if (!(k > 0 && k < nxvz - 1)) continue;
            const int c1 = k + l * nx, c2 = k - 1 + (l - 1) * ncx;
            const int iVzS = c3 - nxvz, iVzW = c3 - 1, iVzE = c3 + 1, iVzN = c3 + nxvz;
            const int iVxNW = c1 + nx - 1, iVxNE = c1 + nx, iVxSW = c1 - 1, iVxSE = c1;
            const int iPrS = c2, iPrN = c2 + ncx, ixyE = c1, ixyW = c1 - 1;
            const double D22S = (l > 0) ? D22[iPrS] : 0.0, D22N = (l < nz - 1) ? D22[iPrN] : 0.0;
            const double D33E = D33[ixyE], D33W = D33[ixyW];
            const double inW = (Gt[ixyW] != 30), inE = (Gt[ixyE] != 30);
            double RegW = 1, RegE = 1, DirW = 0, DirE = 0;
            if (Vt[iVzW] == 13) RegW = 0;
            if (Vt[iVzE] == 13) RegE = 0;
            if (Vt[iVzW] == 11) { DirW = 1; RegW = 0; }
            if (Vt[iVzE] == 11) { DirE = 1; RegE = 0; }
            const double R = D33E * inE * (2 * DirE + RegE) + D33W * inW * (2 * DirW + RegW);
            const double vW = -D33W * RegW * inW / (dx * dx);
            const double vE = -D33E * RegE * inE / (dx * dx);
            const double vS = D22S * (comp - 3) / (3.0 * dz * dz);
            const double vN = D22N * (comp - 3) / (3.0 * dz * dz);
            const double vC = R / (dx * dx) - (D22N + D22S) * (comp - 3) / (3.0 * dz * dz);
            const double uSW = (D22S * comp - 3 * D33W * inW) / (3.0 * dx * dz);
            const double uSE = -(D22S * comp - 3 * D33E * inE) / (3.0 * dx * dz);
            const double uNW = -(D22N * comp - 3 * D33W * inW) / (3.0 * dx * dz);
            const double uNE = (D22N * comp - 3 * D33E * inE) / (3.0 * dx * dz);
            if (Ut[iVxSW] != 30 && l > 0) PUTA(Ut[iVxSW], eqn_u[iVxSW], uSW);
            if (Ut[iVxSE] != 30 && l > 0) PUTA(Ut[iVxSE], eqn_u[iVxSE], uSE);
            if (Ut[iVxNW] != 30 && l < nz - 1) PUTA(Ut[iVxNW], eqn_u[iVxNW], uNW);
            if (Ut[iVxNE] != 30 && l < nz - 1) PUTA(Ut[iVxNE], eqn_u[iVxNE], uNE);
            if (l > 0 && Vt[iVzS] != 30) PUTA(Vt[iVzS], eqn_v[iVzS], vS);
            if (Vt[iVzW] != 30) PUTA(Vt[iVzW], eqn_v[iVzW], vW);
            if (Vt[c3] != 30) PUTA(Vt[c3], eqn_v[c3], vC);
            if (Vt[iVzE] != 30) PUTA(Vt[iVzE], eqn_v[iVzE], vE);
            if (l < nz - 1 && Vt[iVzN] != 30) PUTA(Vt[iVzN], eqn_v[iVzN], vN);
            if (Pt[iPrS] != 30 && l > 0) PUTB(Pt[iPrS], eqn_p[iPrS], -1.0 / dz);
            if (l < nz - 1 && Pt[iPrN] != 30) PUTB(Pt[iPrN], eqn_p[iPrN], 1.0 / dz);
        }
    Ap[row] = na; Bp[row] = nb;
#undef PUTA
#undef PUTB
    /* Continuity rows (divergence), StokesAssemblyDecoupled.c:1108-1127. */
    int nc = 0, prow = 0;
    for (int l = 0; l < ncz; l++)
        for (int k = 0; k < ncx; k++) {
            const int c2 = k + l * ncx;
            if (Pt[c2] != -1) continue;
            Cp[prow++] = nc;
            const int c1 = k + (l + 1) * nx, c3 = k + l * nxvz + 1;
            const int ids[4] = {c1, c1 + 1, c3, c3 + nxvz};
            const double cf[4] = {-1.0 / dx, 1.0 / dx, -1.0 / dz, 1.0 / dz};
            for (int q = 0; q < 4; q++) {
                const signed char t = (q < 2) ? Ut[ids[q]] : Vt[ids[q]];
                if (t == 13 || folds_to_rhs(t)) continue;
                Cj[nc] = (q < 2) ? eqn_u[ids[q]] : eqn_v[ids[q]];
                Cx[nc] = cf[q] * celvol;
                nc++;
            }
        }
    Cp[prow] = nc;
    (void)neq_cont;
}
