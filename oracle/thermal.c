/*
 * Oracle: the implicit energy step of EnergyDirectSolve
 * (ThermalRoutines.c:74-470) with the Jacobi-preconditioned CG of
 * SolveThermalPCG (ThermalSolver.c:338-480, model.thermal_solver = 1), for a
 * Cartesian box without air cells.  TEST INFRASTRUCTURE ONLY.
 *
 * Cells (Ncx x Ncz, all of them unknowns), conductivities kx on the Vx nodes
 * (Nx x (Nz+1)) and kz on the Vz nodes ((Nx+1) x Nz), boundary types and
 * values on the extended cell array BCT_exp ((Ncx+2) x (Ncz+2)): 1 Dirichlet
 * at the wall face through the ghost, 0 insulated (the south ghost adds the
 * reference's flux term -val/dx).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdoracle.h"

/* Assembles the energy matrix (:185-390, theta = 1, transient = 1) as CSR
 * (5 entries per row at most) and its boundary right-hand side bbc; rows in
 * cell order (EvalNumEqTemp numbers every non-air cell in that order). */
void orc_energy_matrix(const orc_grid *g, double dt, const double *rho_n, const double *Cp, const double *kx,
                       const double *kz, const signed char *BCT_t, const double *BCT_v, int *Ap, int *Aj, double *Ax,
                       double *bbc) {
    const int nx = g->Nx, ncx = nx - 1, ncz = g->Nz - 1, nxvz = nx + 1;
    const double idx2 = 1.0 / (g->dx * g->dx), idz2 = 1.0 / (g->dz * g->dz);
    int nnz = 0;
    for (int l = 0; l < ncz; l++)
        for (int k = 0; k < ncx; k++) {
            const int c0 = (k + 1) + (l + 1) * (ncx + 2), c1 = k + (l + 1) * nx, c2 = k + l * ncx, c3 = k + l * nxvz + 1;
            const double rhoCp = rho_n[c2] * Cp[c2];
            const double AW = kx[c1], AE = kx[c1 + 1], AS = kz[c3], AN = kz[c3 + nxvz];
            Ap[c2] = nnz;
            bbc[c2] = 0.0;
            if (l > 0) { Aj[nnz] = c2 - ncx; Ax[nnz++] = -AS * idz2; }
            if (k > 0) { Aj[nnz] = c2 - 1; Ax[nnz++] = -AW * idx2; }
            double val = rhoCp / dt;
            if (l > 0) val += AS * idz2;
            if (k > 0) val += AW * idx2;
            if (k < ncx - 1) val += AE * idx2;
            if (l < ncz - 1) val += AN * idz2;
            if (l == 0 && BCT_t[c0 - (ncx + 2)] == 1) {
                val += 2.0 * AS * idz2;
                bbc[c2] += 2.0 * AS * idz2 * BCT_v[c0 - (ncx + 2)];
            }
            if (l == 0 && BCT_t[c0 - (ncx + 2)] == 0) bbc[c2] += -1.0 / g->dx * BCT_v[c0 - (ncx + 2)];
            if (k == 0 && BCT_t[c0 - 1] == 1) {
                val += 2.0 * AW * idx2;
                bbc[c2] += 2.0 * AW * idx2 * BCT_v[c0 - 1];
            }
            if (k == ncx - 1 && BCT_t[c0 + 1] == 1) {
                val += 2.0 * AE * idx2;
                bbc[c2] += 2.0 * AE * idx2 * BCT_v[c0 + 1];
            }
            if (l == ncz - 1 && BCT_t[c0 + (ncx + 2)] == 1) {
                val += 2.0 * AN * idz2;
                bbc[c2] += 2.0 * AN * idz2 * BCT_v[c0 + (ncx + 2)];
            }
            Aj[nnz] = c2; Ax[nnz++] = val;
            if (k < ncx - 1) { Aj[nnz] = c2 + 1; Ax[nnz++] = -AE * idx2; }
            if (l < ncz - 1) { Aj[nnz] = c2 + ncx; Ax[nnz++] = -AN * idz2; }
        }
    Ap[ncx * ncz] = nnz;
}

/* One energy step: b = rho Cp T / dt + Qr (+ Wdiss) (+ Ha) + bbc in the
 * reference's order (:140-165, :395), then Jacobi PCG from x0 = T with the
 * reference's stopping test ||r|| / ||r0|| < rel_tol (:400-470 of
 * ThermalSolver.c), T <- x.  Returns the iteration count (0: the start
 * already solves it; negative: the reference's breakdown or no-convergence
 * codes, -1 after max_its). */
int orc_energy_step(const orc_grid *g, double dt, const double *rho_n, const double *Cp, const double *Qr,
                    const double *Wdiss, const double *Ha, const double *kx, const double *kz, const signed char *BCT_t,
                    const double *BCT_v, int max_its, double rel_tol, double *T) {
    const int n = (g->Nx - 1) * (g->Nz - 1);
    int *Ap = (int *)malloc(sizeof(int) * (n + 1)), *Aj = (int *)malloc(sizeof(int) * 5 * (size_t)n);
    double *Ax = (double *)malloc(sizeof(double) * 5 * (size_t)n);
    double *w = (double *)calloc(7 * (size_t)n, sizeof(double));
    double *bbc = w, *b = w + n, *r = w + 2 * n, *z = w + 3 * n, *p = w + 4 * n, *q = w + 5 * n, *Minv = w + 6 * n;
    orc_energy_matrix(g, dt, rho_n, Cp, kx, kz, BCT_t, BCT_v, Ap, Aj, Ax, bbc);
    for (int i = 0; i < n; i++) {
        b[i] = rho_n[i] * Cp[i] * T[i] / dt;
        b[i] += 1.0 * (Qr ? Qr[i] : 0.0);
        if (Wdiss) b[i] += Wdiss[i];
        if (Ha) b[i] += Ha[i];
        b[i] += 1.0 * bbc[i];
    }
    for (int i = 0; i < n; i++) {
        double d = 0.0;
        for (int k = Ap[i]; k < Ap[i + 1]; k++)
            if (Aj[k] == i) { d = Ax[k]; break; }
        Minv[i] = (fabs(d) > 1e-30) ? 1.0 / d : 1.0;
    }
    double *x = T;
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int k = Ap[i]; k < Ap[i + 1]; k++) s += Ax[k] * x[Aj[k]];
        r[i] = b[i] - s;
    }
    double rz = 0.0, r0 = 0.0;
    for (int i = 0; i < n; i++) { z[i] = Minv[i] * r[i]; p[i] = z[i]; }
    for (int i = 0; i < n; i++) rz += r[i] * z[i];
    for (int i = 0; i < n; i++) r0 += r[i] * r[i];
    r0 = sqrt(r0);
    int its = -1;
    if (r0 < 1e-30) its = 0;
    for (int it = 0; its < 0 && it < max_its; it++) {
        for (int i = 0; i < n; i++) {
            double s = 0.0;
            for (int k = Ap[i]; k < Ap[i + 1]; k++) s += Ax[k] * p[Aj[k]];
            q[i] = s;
        }
        double pq = 0.0;
        for (int i = 0; i < n; i++) pq += p[i] * q[i];
        if (fabs(pq) < 1e-30) { its = -2; break; }
        const double alpha = rz / pq;
        for (int i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
        double rn = 0.0;
        for (int i = 0; i < n; i++) rn += r[i] * r[i];
        rn = sqrt(rn);
        if (rn / r0 < rel_tol) { its = it + 1; break; }
        for (int i = 0; i < n; i++) z[i] = Minv[i] * r[i];
        double rzn = 0.0;
        for (int i = 0; i < n; i++) rzn += r[i] * z[i];
        if (fabs(rz) < 1e-30) { its = -3; break; }
        const double beta = rzn / rz;
        rz = rzn;
        for (int i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
    }
    free(Ap); free(Aj); free(Ax); free(w);
    return its;
}
