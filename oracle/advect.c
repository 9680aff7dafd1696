/*
 * Marker advection: V2P (ParticleRoutines.c:63-256) and the fourth-order
 * Runge-Kutta step of RogerGuntherII (AdvectionRoutines.c:347-410) with
 * isoutPart (AdvectionRoutines.c:320-341, non-periodic box).
 * TEST INFRASTRUCTURE ONLY (see mdoracle.h).
 */
#include <math.h>
#include <stdlib.h>

#include "mdoracle.h"

/* the cell of a staggered component: ceil(d/h) - 1 clamped to [0, n-2]
 * (ParticleRoutines.c:81-88, 92-99) */
static int v2p_cell(double d, double h, int n) {
    int j = (int)ceil(d / h) - 1;
    if (j < 0) j = 0;
    if (j > n - 2) j = n - 2;
    return j;
}

/* Taras' conservative correction of one interpolated row/column
 * (ParticleRoutines.c:127-158, 216-248): 0.5 (t - 0.5)^2 times the second
 * difference on the side of the marker. */
static double v2p_corr(double t) { return 0.5 * pow(t - 0.5, 2.0); }

/* V2P for one active marker (ParticleRoutines.c:76-252). u: Vx on
 * Nx x (Nz+1) nodes (xg, zvx); v: Vz on (Nx+1) x Nz nodes (xvz, zg). */
static void v2p_one(const orc_grid *g, const double *u, const double *v, int conserv, double x, double z,
                    double *vxm, double *vzm) {
    const double dx = g->dx, dz = g->dz;
    const signed char *tu = g->BCu_t, *tv = g->BCv_t;
    /* Vx */
    {
        const int Nx = g->Nx, Nz = g->Nz + 1;
        const int j = v2p_cell(x - g->xg[0], dx, Nx), i = v2p_cell(z - g->zvx[0], dz, Nz);
        const double dxm = fabs(x - g->xg[j]), dzm = fabs(z - g->zvx[i]);
        const int iSW = j + i * Nx, iSE = iSW + 1, iNW = j + (i + 1) * Nx, iNE = iNW + 1;
        double VNW = u[iNW], VNE = u[iNE], VSW = u[iSW], VSE = u[iSE];
        if (tu[iSW] == 30) VSW = u[iNW];
        if (tu[iSE] == 30) VSE = u[iNE];
        if (tu[iNW] == 30) VNW = u[iSW];
        if (tu[iNE] == 30) VNE = u[iSE];
        if (tu[iNE] == 30 && tu[iSE] == 30) { VNE = VSW; VSE = VSW; }
        if (tu[iNW] == 30 && tu[iSW] == 30) { VNW = VSE; VSW = VSE; }
        double v13 = VSW * (1.0 - dxm / dx) + VSE * dxm / dx;
        double v24 = VNW * (1.0 - dxm / dx) + VNE * dxm / dx;
        if (conserv == 1) {
            if (dxm / dx >= 0.5) {
                if (j < Nx - 2) {
                    const int iSEE = iSE + 1, iNEE = iNE + 1;
                    double VSEE = u[iSEE], VNEE = u[iNEE];
                    if (tu[iSEE] == 30) VSEE = u[iSW];
                    if (tu[iNEE] == 30) VNEE = u[iNW];
                    v13 = v13 + v2p_corr(dxm / dx) * (VSW - 2.0 * VSE + VSEE);
                    v24 = v24 + v2p_corr(dxm / dx) * (VNW - 2.0 * VNE + VNEE);
                }
            } else if (j > 0) {
                const int iSWW = iSW - 1, iNWW = iNW - 1;
                double VSWW = u[iSWW], VNWW = u[iNWW];
                if (tu[iSWW] == 30) VSWW = u[iSE];
                if (tu[iNWW] == 30) VNWW = u[iNE];
                v13 = v13 + v2p_corr(dxm / dx) * (VSWW - 2.0 * VSW + VSE);
                v24 = v24 + v2p_corr(dxm / dx) * (VNWW - 2.0 * VNW + VNE);
            }
        }
        *vxm = (1.0 - dzm / dz) * v13 + (dzm / dz) * v24;
    }
    /* Vz */
    {
        const int Nx = g->Nx + 1, Nz = g->Nz;
        const int j = v2p_cell(x - g->xvz[0], dx, Nx), i = v2p_cell(z - g->zg[0], dz, Nz);
        const double dxm = fabs(x - g->xvz[j]), dzm = fabs(z - g->zg[i]);
        const int iSW = j + i * Nx, iSE = iSW + 1, iNW = j + (i + 1) * Nx, iNE = iNW + 1;
        double VSE = v[iSE], VSW = v[iSW], VNE = v[iNE], VNW = v[iNW];
        if (tv[iSW] == 30) VSW = v[iNE];
        if (tv[iSE] == 30) VSE = v[iNE];
        if (tv[iNW] == 30) VNW = v[iSW];
        if (tv[iNE] == 30) VNE = v[iSE];
        if (tv[iNE] == 30 && tv[iSE] == 30) { VNE = VSW; VSE = VSW; }
        if (tv[iNW] == 30 && tv[iSW] == 30) { VNW = VSE; VSW = VSE; }
        double v12 = VSW * (1.0 - dzm / dz) + VNW * dzm / dz;
        double v34 = VSE * (1.0 - dzm / dz) + VNE * dzm / dz;
        if (conserv == 1) {
            if (dzm / dz >= 0.5) {
                if (i < Nz - 2) {
                    const int iNNW = j + (i + 2) * Nx, iNNE = iNNW + 1;
                    double VNNW = v[iNNW], VNNE = v[iNNE];
                    if (tv[iNNW] == 30) VNNW = v[iSW];
                    if (tv[iNNE] == 30) VNNE = v[iSE];
                    v12 = v12 + v2p_corr(dzm / dz) * (VSW - 2.0 * VNW + VNNW);
                    v34 = v34 + v2p_corr(dzm / dz) * (VSE - 2.0 * VNE + VNNE);
                }
            } else if (i > 0) {
                const int iSSW = j + (i - 1) * Nx, iSSE = iSSW + 1;
                double VSSW = v[iSSW], VSSE = v[iSSE];
                if (tv[iSSW] == 30) VSSW = v[iNW];
                if (tv[iSSE] == 30) VSSE = v[iNE];
                v12 = v12 + v2p_corr(dzm / dz) * (VSSW - 2.0 * VSW + VNW);
                v34 = v34 + v2p_corr(dzm / dz) * (VSSE - 2.0 * VSE + VNE);
            }
        }
        *vzm = (1.0 - dxm / dx) * v12 + (dxm / dx) * v34;
    }
}

void orc_v2p(const orc_grid *g, const double *u, const double *v, int conserv, int n, const double *x,
             const double *z, const int *phase, double *vx, double *vz) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        vx[k] = 0.0;
        vz[k] = 0.0;
        if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &vx[k], &vz[k]);
    }
}

/* isoutPart, non-periodic (AdvectionRoutines.c:336-338) */
static void isout(double x, double z, const double box[4], int *phase) {
    if (x < box[0] || x > box[1] || z < box[2] || z > box[3]) *phase = -1;
}

/* RogerGuntherII's loop body (AdvectionRoutines.c:372-405) for every marker */
void orc_advect_rk4(const orc_grid *g, const double *u, const double *v, double dt, int conserv,
                    double xmin, double xmax, double zmin, double zmax, int n, double *x, double *z,
                    int *phase, double *vx, double *vz) {
    const double box[4] = {xmin, xmax, zmin, zmax};
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        const double xA = x[k], zA = z[k];
        double VxA = 0.0, VzA = 0.0, VxB = 0.0, VzB = 0.0, VxC = 0.0, VzC = 0.0, VxD = 0.0, VzD = 0.0;
        if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &VxA, &VzA);
        if (phase[k] != -1) {
            x[k] = xA + 0.5 * dt * VxA;
            z[k] = zA + 0.5 * dt * VzA;
        }
        isout(x[k], z[k], box, &phase[k]);
        if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &VxB, &VzB);
        if (phase[k] != -1) {
            x[k] = xA + 0.5 * dt * VxB;
            z[k] = zA + 0.5 * dt * VzB;
        }
        isout(x[k], z[k], box, &phase[k]);
        if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &VxC, &VzC);
        if (phase[k] != -1) {
            x[k] = xA + 1.0 * dt * VxC;
            z[k] = zA + 1.0 * dt * VzC;
        }
        isout(x[k], z[k], box, &phase[k]);
        if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &VxD, &VzD);
        VxA = (1.0 / 6.0) * (VxA + 2.0 * VxB + 2.0 * VxC + VxD);
        VzA = (1.0 / 6.0) * (VzA + 2.0 * VzB + 2.0 * VzC + VzD);
        if (phase[k] != -1) {
            x[k] = xA + dt * VxA;
            z[k] = zA + dt * VzA;
            if (vx) vx[k] = VxA;
            if (vz) vz[k] = VzA;
        }
        isout(x[k], z[k], box, &phase[k]);
    }
}

/* Interp_Grid2P_centroids2 (ParticleRoutines.c:1332-1580), non-periodic:
 * a centre field to the markers, bilinear between the centres around the
 * marker (extended coordinates xvz, zvx), values copied along the box edges,
 * air / surface cells (BCp tag 30/31) replaced by their vertical neighbour,
 * normalised by the weight sum. Inactive markers get 0. */
static double g2p_cent_one(const orc_grid *g, const double *F, double x, double z) {
    const int Nx = g->Nx - 1, Nz = g->Nz - 1;
    const double dx = g->dx, dz = g->dz;
    const double *X = g->xvz, *Z = g->zvx;
    const signed char *tag = g->BCp_t;
    const int j = (int)ceil((x - X[0]) / dx) - 1;
    const int i = (int)ceil((z - Z[0]) / dz) - 1;
    int ic = 0, jc = 0, iSW = 0, iSE = 0, iNW = 0, iNE = 0;
    if (j > 0 && j < Nx && i > 0 && i < Nz) {
        ic = i - 1; jc = j - 1;
        iSW = jc + ic * Nx; iSE = iSW + 1; iNW = jc + (ic + 1) * Nx; iNE = iNW + 1;
    } else if (j == 0 && i > 0 && i < Nz) {
        ic = i - 1; jc = j;
        iSW = iSE = jc + ic * Nx; iNW = iNE = jc + (ic + 1) * Nx;
    } else if (j == Nx && i > 0 && i < Nz) {
        ic = i - 1; jc = j - 1;
        iSW = iSE = jc + ic * Nx; iNW = iNE = jc + (ic + 1) * Nx;
    } else if (i == 0 && j > 0 && j < Nx) {
        ic = i; jc = j - 1;
        iSW = iNW = jc + ic * Nx; iSE = iNE = jc + ic * Nx + 1;
    } else if (i == Nz && j > 0 && j < Nx) {
        ic = i - 1; jc = j - 1;
        iSW = iNW = jc + ic * Nx; iSE = iNE = jc + ic * Nx + 1;
    } else if (i == 0 && j == 0) {
        iSW = iSE = iNW = iNE = 0;
    } else if (i == 0 && j == Nx) {
        iSW = iSE = iNW = iNE = j - 1;
    } else if (i == Nz && j == 0) {
        iSW = iSE = iNW = iNE = (i - 1) * Nx;
    } else if (i == Nz && j == Nx) {
        iSW = iSE = iNW = iNE = (j - 1) + (i - 1) * Nx;
    }
    (void)ic; (void)jc;
    /* markers of the box always take one of the cases above; other inputs
     * read clamped coordinates (the reference leaves them undefined) */
    const int jj = j < 0 ? 0 : (j > Nx ? Nx : j), ii = i < 0 ? 0 : (i > Nz ? Nz : i);
    const double dxm = x - X[jj], dzm = z - Z[ii];
    double val = 0.0, sumW = 0.0;
    const int air_SW = tag[iSW] == 30 || tag[iSW] == 31, air_SE = tag[iSE] == 30 || tag[iSE] == 31;
    const int air_NW = tag[iNW] == 30 || tag[iNW] == 31, air_NE = tag[iNE] == 30 || tag[iNE] == 31;
    val += (1.0 - dxm / dx) * (1.0 - dzm / dz) * F[air_SW ? iNW : iSW];
    sumW += (1.0 - dxm / dx) * (1.0 - dzm / dz);
    val += (dxm / dx) * (1.0 - dzm / dz) * F[air_SE ? iNE : iSE];
    sumW += (dxm / dx) * (1.0 - dzm / dz);
    val += (1.0 - dxm / dx) * (dzm / dz) * F[air_NW ? iSW : iNW];
    sumW += (1.0 - dxm / dx) * (dzm / dz);
    val += (dxm / dx) * (dzm / dz) * F[air_NE ? iSE : iNE];
    sumW += (dxm / dx) * (dzm / dz);
    if (sumW > 1e-13) val /= sumW;
    return val;
}

void orc_g2p_centroids(const orc_grid *g, const double *F, int n, const double *x, const double *z,
                       const int *phase, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) out[k] = (phase[k] != -1) ? g2p_cent_one(g, F, x[k], z[k]) : 0.0;
}

/* UpdateParticlePressure (RheologyParticles.c:1621-1704) without subgrid
 * diffusion (model.subgrid_diffusion < 2): dp = p_in - p0_n on the cells
 * that are not air (:1631-1637), interpolated to the markers and added to
 * their pressure (:1694-1700). dp (ncells) receives the increment. */
void orc_update_particle_pressure(const orc_grid *g, const double *p_in, const double *p0_n, double *dp,
                                  int n, const double *x, const double *z, const int *phase, double *P) {
    const int nc = (g->Nx - 1) * (g->Nz - 1);
    for (int c = 0; c < nc; c++) dp[c] = (g->BCp_t[c] != 30) ? (p_in[c] - p0_n[c]) : 0.0;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        const double inc = (phase[k] != -1) ? g2p_cent_one(g, dp, x[k], z[k]) : 0.0;
        P[k] = P[k] + inc;
    }
}

/* Interp_Grid2P (ParticleRoutines.c:1223-1325) on the vertices (xg, zg,
 * BCg tags): bilinear in the clamped cell, tag 30/31 vertices replaced by
 * their vertical neighbour, no normalisation (the reference sums the
 * weights but does not divide); inactive markers get 0. */
static double g2p_vert_one(const orc_grid *g, const double *F, double x, double z) {
    const int Nx = g->Nx, Nz = g->Nz;
    const double dx = g->dx, dz = g->dz;
    const signed char *tag = g->BCg_t;
    int j = (int)ceil((x - g->xg[0]) / dx) - 1;
    if (j < 0) j = 0;
    if (j > Nx - 2) j = Nx - 2;
    int i = (int)ceil((z - g->zg[0]) / dz) - 1;
    if (i < 0) i = 0;
    if (i > Nz - 2) i = Nz - 2;
    const double dxm = fabs(x - g->xg[j]), dzm = fabs(z - g->zg[i]);
    const int iSW = j + i * Nx, iSE = iSW + 1, iNW = j + (i + 1) * Nx, iNE = iNW + 1;
    const int aSW = tag[iSW] == 30 || tag[iSW] == 31, aSE = tag[iSE] == 30 || tag[iSE] == 31;
    const int aNW = tag[iNW] == 30 || tag[iNW] == 31, aNE = tag[iNE] == 30 || tag[iNE] == 31;
    double val = 0.0;
    val += (1.0 - dxm / dx) * (1.0 - dzm / dz) * F[aSW ? iNW : iSW];
    val += (dxm / dx) * (1.0 - dzm / dz) * F[aSE ? iNE : iSE];
    val += (1.0 - dxm / dx) * (dzm / dz) * F[aNW ? iSW : iNW];
    val += (dxm / dx) * (dzm / dz) * F[aNE ? iSE : iNE];
    return val;
}

void orc_g2p_vertices(const orc_grid *g, const double *F, int n, const double *x, const double *z,
                      const int *phase, double *out) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) out[k] = (phase[k] != -1) ? g2p_vert_one(g, F, x[k], z[k]) : 0.0;
}

/* UpdateParticleStress (RheologyParticles.c:1725-2082) with elastic = 1,
 * subgrid_diffusion 0 or 1 and no stress rotation (:1924-1960): the grid
 * stress increments (sxxd - sxxd0, szzd - szzd0 on the cells not tagged
 * 30/31, sxz - sxz0 on the vertices not tagged 30) interpolated to the
 * markers and added to their stresses. */
void orc_update_particle_stress(const orc_grid *g, const double *sxxd, const double *szzd, const double *sxz,
                                const double *sxxd0, const double *szzd0, const double *sxz0, int n,
                                const double *x, const double *z, const int *phase, double *msxxd,
                                double *mszzd, double *msxz) {
    const int nc = (g->Nx - 1) * (g->Nz - 1), nv = g->Nx * g->Nz;
    double *dxx = calloc((size_t)nc, sizeof(double)), *dzz = calloc((size_t)nc, sizeof(double));
    double *dxz = calloc((size_t)nv, sizeof(double));
    for (int c = 0; c < nc; c++) {
        if (g->BCp_t[c] != 30 && g->BCp_t[c] != 31) dxx[c] = sxxd[c] - sxxd0[c];
        if (g->BCp_t[c] != 30 && g->BCp_t[c] != 31) dzz[c] = szzd[c] - szzd0[c];
    }
    for (int c = 0; c < nv; c++)
        if (g->BCg_t[c] != 30) dxz[c] = sxz[c] - sxz0[c];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        const int on = phase[k] != -1;
        msxxd[k] = msxxd[k] + (on ? g2p_cent_one(g, dxx, x[k], z[k]) : 0.0);
        mszzd[k] = mszzd[k] + (on ? g2p_cent_one(g, dzz, x[k], z[k]) : 0.0);
        msxz[k] = msxz[k] + (on ? g2p_vert_one(g, dxz, x[k], z[k]) : 0.0);
    }
    free(dxx);
    free(dzz);
    free(dxz);
}

/* AccumulatedStrainII (RheologyParticles.c:346-649): per-centre increments
 * dt * eII (the total being the sum of the seven mechanisms in the
 * reference's order), then per active marker the bilinear interpolation of
 * each increment from the four centres around it (centre coordinates xc,
 * zc; the cell index clamped to the centre grid; BCp tag 30/31 centres left
 * out and the weights renormalised when their sum exceeds 1e-13), added to
 * the marker's accumulated strains.  eII[0..6] = pl, pwl, exp, lin, gbs, el,
 * cst (NULL = 0); strain[0..7] = total, el, pl, pl_vol, pwl, exp, lin, gbs
 * (NULL = skipped).  Returns the first cell with a negative total increment
 * (the reference exits there), or -1. */
int orc_accumulated_strain(const orc_grid *g, double dt, const double *const *eII, const double *div_u_pl, int n,
                           const double *x, const double *z, const int *phase, double *const *strain) {
    const int Nx = g->Nx - 1, Nz = g->Nz - 1, nc = Nx * Nz;   /* centre grid */
    const double dx = g->dx, dz = g->dz;
    double *inc = (double *)malloc(sizeof(double) * 8 * (size_t)nc);
    int bad = -1;
    for (int c = 0; c < nc; c++) {
        double e[7];
        for (int q = 0; q < 7; q++) e[q] = eII[q] ? eII[q][c] : 0.0;
        inc[c] = dt * (e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6]);
        if (inc[c] < 0 && bad < 0) bad = c;
        inc[1 * nc + c] = dt * e[5];                                   /* elastic    */
        inc[2 * nc + c] = dt * e[0];                                   /* plastic    */
        inc[3 * nc + c] = dt * (div_u_pl ? div_u_pl[c] : 0.0);         /* pl. volume */
        inc[4 * nc + c] = dt * e[1];                                   /* power law  */
        inc[5 * nc + c] = dt * e[2];                                   /* exponential*/
        inc[6 * nc + c] = dt * e[3];                                   /* linear     */
        inc[7 * nc + c] = dt * e[4];                                   /* GBS        */
    }
    if (bad >= 0) { free(inc); return bad; }
    for (int k = 0; k < n; k++) {
        double dE[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (phase[k] != -1) {
            int j = (int)ceil((x[k] - g->xc[0]) / dx) - 1;
            if (j < 0) j = 0;
            if (j > Nx - 2) j = Nx - 2;
            int i = (int)ceil((z[k] - g->zc[0]) / dz) - 1;
            if (i < 0) i = 0;
            if (i > Nz - 2) i = Nz - 2;
            const double dxm = x[k] - g->xc[j], dzm = z[k] - g->zc[i];
            const int nd[4] = {j + i * Nx, j + i * Nx + 1, j + (i + 1) * Nx, j + (i + 1) * Nx + 1};
            const double w[4] = {(1.0 - dxm / dx) * (1.0 - dzm / dz), (dxm / dx) * (1.0 - dzm / dz),
                                 (1.0 - dxm / dx) * (dzm / dz), (dxm / dx) * (dzm / dz)};
            double sumW = 0.0;
            for (int q = 0; q < 4; q++) {
                const signed char t = g->BCp_t[nd[q]];
                if (t == 30 || t == 31) continue;
                for (int f = 0; f < 8; f++) dE[f] += w[q] * inc[(size_t)f * nc + nd[q]];
                sumW += w[q];
            }
            if (sumW > 1e-13)
                for (int f = 0; f < 8; f++) dE[f] /= sumW;
        }
        for (int f = 0; f < 8; f++)
            if (strain[f]) strain[f][k] += dE[f];
    }
    free(inc);
    return -1;
}

/* Stress rotation of UpdateParticleStress (RheologyParticles.c:1965-1988) on
 * the active markers with their spin wxzm: mode 1 the Jaumann rate, mode 2
 * the rotation by the angle dt*wxzm. */
void orc_rotate_stress(int mode, double dt, int n, const int *phase, const double *wxzm, double *sxxd, double *szzd,
                       double *sxz) {
    for (int k = 0; k < n; k++) {
        if (phase[k] == -1) continue;
        const double txx = sxxd[k], tzz = szzd[k], txz = sxz[k];
        if (mode == 1) {
            sxxd[k] -= dt * 2.0 * txz * wxzm[k];
            szzd[k] -= -dt * 2.0 * txz * wxzm[k];
            sxz[k] -= dt * (tzz - txx) * wxzm[k];
        } else if (mode == 2) {
            const double a = dt * wxzm[k], c = cos(a), s = sin(a);
            sxxd[k] = (txx * c + txz * s) * c + (txz * c + tzz * s) * s;
            szzd[k] = -(-txx * s + txz * c) * s + (-txz * s + tzz * c) * c;
            sxz[k] = -(txx * c + txz * s) * s + (txz * c + tzz * s) * c;
        }
    }
}

/* UpdateParticleStress with subgrid_diffusion = 2 (RheologyParticles.c:
 * 1811-1895): old grid stresses and the vertex viscosity eta_phys_s at the
 * markers; per active marker the subgrid relaxation
 *   dtau_s = -(tau_m - tau0_m) (1 - exp(-d dt / (eta_m / G_p))),  d = 1;
 * the subgrid increments back to the grid (P2Mastah, flag 1, arithmetic,
 * with the BCp / BCg node tags), the remaining grid increments
 * (tau - tau0) - dtau_s_grid on non-air nodes, those at the markers, and
 * tau_m += dtau_s + dtau_r. */
void orc_update_particle_stress_subgrid(const orc_grid *g, double dt, const double *G_phase, int stencil,
                                        const double *sxxd, const double *szzd, const double *sxz,
                                        const double *sxxd0, const double *szzd0, const double *sxz0,
                                        const double *eta_phys_s, int n, const double *x, const double *z,
                                        const int *phase, double *msxxd, double *mszzd, double *msxz) {
    const int nc = (g->Nx - 1) * (g->Nz - 1), nv = g->Nx * g->Nz;
    const size_t N = (size_t)(n > 0 ? n : 1);
    double *m0[3], *ms[3], *mr[3], *etam = calloc(N, sizeof(double));
    for (int q = 0; q < 3; q++) {
        m0[q] = calloc(N, sizeof(double));
        ms[q] = calloc(N, sizeof(double));
        mr[q] = calloc(N, sizeof(double));
    }
    orc_g2p_centroids(g, sxxd0, n, x, z, phase, m0[0]);
    orc_g2p_centroids(g, szzd0, n, x, z, phase, m0[1]);
    orc_g2p_vertices(g, sxz0, n, x, z, phase, m0[2]);
    orc_g2p_vertices(g, eta_phys_s, n, x, z, phase, etam);
    double *tm[3] = {msxxd, mszzd, msxz};
    const double d = 1.0;
    for (int k = 0; k < n; k++) {
        if (phase[k] == -1) continue;
        const double dtaum = etam[k] / G_phase[phase[k]];
        for (int q = 0; q < 3; q++) ms[q][k] = -(tm[q][k] - m0[q][k]) * (1.0 - exp(-d * dt / dtaum));
    }
    double *gs[3] = {calloc((size_t)nc, sizeof(double)), calloc((size_t)nc, sizeof(double)),
                     calloc((size_t)nv, sizeof(double))};
    double *gr[3] = {calloc((size_t)nc, sizeof(double)), calloc((size_t)nc, sizeof(double)),
                     calloc((size_t)nv, sizeof(double))};
    for (int q = 0; q < 3; q++)
        orc_p2mastah(g, n, x, z, phase, ms[q], 1, 0, 0, q < 2 ? 1 : 0, stencil, 0, q < 2 ? g->BCp_t : g->BCg_t,
                     gs[q], 0, NULL, 0, 1);
    for (int c = 0; c < nc; c++) {
        if (g->BCp_t[c] != 30 && g->BCp_t[c] != 31) gr[0][c] = (sxxd[c] - sxxd0[c]) - gs[0][c];
        if (g->BCp_t[c] != 30 && g->BCp_t[c] != 31) gr[1][c] = (szzd[c] - szzd0[c]) - gs[1][c];
    }
    for (int c = 0; c < nv; c++)
        if (g->BCg_t[c] != 30) gr[2][c] = (sxz[c] - sxz0[c]) - gs[2][c];
    orc_g2p_centroids(g, gr[0], n, x, z, phase, mr[0]);
    orc_g2p_centroids(g, gr[1], n, x, z, phase, mr[1]);
    orc_g2p_vertices(g, gr[2], n, x, z, phase, mr[2]);
    for (int k = 0; k < n; k++)
        if (phase[k] != -1)
            for (int q = 0; q < 3; q++) tm[q][k] = tm[q][k] + ms[q][k] + mr[q][k];
    for (int q = 0; q < 3; q++) { free(m0[q]); free(ms[q]); free(mr[q]); free(gs[q]); free(gr[q]); }
    free(etam);
}

/* UpdateParticlePressure with subgrid_diffusion >= 2 (RheologyParticles.c:
 * 1621-1695): dp = p_in - p0_n on non-air cells; eta_phys_n and p0_n (BCt
 * 30 cells left at 0; here the BCp tags, which mark the same air cells) at
 * the markers; dP_s = -(P_m - P0_m) (1 - exp(-d dt / (eta_m / beta_p))) (the
 * reference's relaxation time, kept as written); P2Mastah of dP_s onto the
 * centres (stencil 1, as the reference hard-codes), remaining dp - dP_s_grid
 * back to the markers, P_m += dP_s + dP_r. */
void orc_update_particle_pressure_subgrid(const orc_grid *g, double dt, const double *beta_phase,
                                          const double *p_in, const double *p0_n, const double *eta_phys_n,
                                          double *dp, int n, const double *x, const double *z, const int *phase,
                                          double *P) {
    const int nc = (g->Nx - 1) * (g->Nz - 1);
    const size_t N = (size_t)(n > 0 ? n : 1);
    for (int c = 0; c < nc; c++) dp[c] = (g->BCp_t[c] != 30) ? (p_in[c] - p0_n[c]) : 0.0;
    double *etam = calloc(N, sizeof(double)), *Pm0 = calloc(N, sizeof(double)), *dPms = calloc(N, sizeof(double)),
           *dPmr = calloc(N, sizeof(double));
    double *Pg0 = calloc((size_t)nc, sizeof(double)), *dPgs = calloc((size_t)nc, sizeof(double)),
           *dPgr = calloc((size_t)nc, sizeof(double));
    orc_g2p_centroids(g, eta_phys_n, n, x, z, phase, etam);
    for (int c = 0; c < nc; c++)
        if (g->BCp_t[c] != 30) Pg0[c] = p0_n[c];
    orc_g2p_centroids(g, Pg0, n, x, z, phase, Pm0);
    const double d = 1.0;
    for (int k = 0; k < n; k++) {
        if (phase[k] == -1) continue;
        const double dtm = etam[k] / (beta_phase[phase[k]]);
        dPms[k] = -(P[k] - Pm0[k]) * (1.0 - exp(-d * dt / dtm));
    }
    orc_p2mastah(g, n, x, z, phase, dPms, 1, 0, 0, 1, 1, 0, g->BCp_t, dPgs, 0, NULL, 0, 1);
    for (int c = 0; c < nc; c++)
        if (g->BCp_t[c] != 30) dPgr[c] = dp[c] - dPgs[c];
    orc_g2p_centroids(g, dPgr, n, x, z, phase, dPmr);
    for (int k = 0; k < n; k++)
        if (phase[k] != -1) P[k] = P[k] + dPms[k] + dPmr[k];
    free(etam); free(Pm0); free(dPms); free(dPmr); free(Pg0); free(dPgs); free(dPgr);
}
