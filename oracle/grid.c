/*
 * Oracle: grid coordinates, boundary tags, DOF numbering, marker index,
 * P2Mastah, centroid->vertex averaging, ghost BCs, strain rates, body force.
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "mdoracle.h"

/* GridRoutines.c:246-330: every coordinate vector is a running sum. */
void orc_grid_coords(int Nx, int Nz, double xmin, double zmin, double xmax, double zmax,
                     double *dx_out, double *dz_out, double *xg, double *zg, double *xc,
                     double *zc, double *xvz, double *zvx) {
    const double dx = (xmax - xmin) / (Nx - 1);
    const double dz = (zmax - zmin) / (Nz - 1);
    *dx_out = dx;
    *dz_out = dz;
    xg[0] = xmin;
    for (int i = 1; i < Nx; i++) xg[i] = xg[i - 1] + dx;
    zg[0] = zmin;
    for (int i = 1; i < Nz; i++) zg[i] = zg[i - 1] + dz;
    xc[0] = xmin + dx / 2.0;
    for (int i = 1; i < Nx - 1; i++) xc[i] = xc[i - 1] + dx;
    zc[0] = zmin + dz / 2.0;
    for (int i = 1; i < Nz - 1; i++) zc[i] = zc[i - 1] + dz;
    xvz[0] = xmin - dx / 2.0;
    for (int i = 1; i < Nx + 1; i++) xvz[i] = xvz[i - 1] + dx;
    zvx[0] = zmin - dz / 2.0;
    for (int i = 1; i < Nz + 1; i++) zvx[i] = zvx[i - 1] + dz;
}

/* Setup.c:222-262 (Vx positions), :443-480 (Vz positions), :568-603 (P),
 * with the pure-shear templates Setup.c:907-924 and :964-981. */
void orc_bc_pure_shear(orc_grid *g, double ebg) {
    const int Nx = g->Nx, Nz = g->Nz;
    for (int l = 0; l < Nz + 1; l++)
        for (int k = 0; k < Nx; k++) {
            const int c = k + l * Nx;
            const int edge_x = (k == 0 || k == Nx - 1);
            const int edge_z = (l == 0 || l == Nz);
            if (edge_x && !edge_z) {          /* W, E: normal velocity */
                g->BCu_t[c] = 0;
                g->BCu_v[c] = -g->xg[k] * ebg;
            } else if (edge_x || edge_z) {    /* corners, S, N ghosts  */
                g->BCu_t[c] = 13;
                g->BCu_v[c] = 0.0;
            } else {
                g->BCu_t[c] = -1;
                g->BCu_v[c] = 0.0;
            }
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx + 1; k++) {
            const int c = k + l * (Nx + 1);
            const int edge_x = (k == 0 || k == Nx);
            const int edge_z = (l == 0 || l == Nz - 1);
            if (edge_x) {                     /* W, E and corners      */
                g->BCv_t[c] = 13;
                g->BCv_v[c] = 0.0;
            } else if (edge_z) {              /* S, N: normal velocity */
                g->BCv_t[c] = 0;
                g->BCv_v[c] = g->zg[l] * ebg;
            } else {
                g->BCv_t[c] = -1;
                g->BCv_v[c] = 0.0;
            }
        }
    for (int c = 0; c < (Nx - 1) * (Nz - 1); c++) g->BCp_t[c] = -1;
    for (int c = 0; c < Nx * Nz; c++) g->BCg_t[c] = -1;
}

/* StokesRoutines.c:1162-1220: Vx rows, then Vz rows, then pressure. */
int orc_eqn_numbering(const orc_grid *g, int *eqn_u, int *eqn_v, int *eqn_p, int *neq_mom,
                      int *neq_cont) {
    const int Nx = g->Nx, Nz = g->Nz;
    int inc = 0, incp = 0;
    for (int l = 0; l < Nz + 1; l++)
        for (int k = 0; k < Nx; k++) {
            const int c = k + l * Nx;
            const signed char t = g->BCu_t[c];
            eqn_u[c] = -1;
            if (t == -1 || t == 2 || t == -2) eqn_u[c] = inc++;
            if (t == -12) eqn_u[c] = eqn_u[c - Nx + 1];
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx + 1; k++) {
            const int c = k + l * (Nx + 1);
            const signed char t = g->BCv_t[c];
            eqn_v[c] = -1;
            if (t == -1 || t == 2) eqn_v[c] = inc++;
        }
    *neq_mom = inc;
    for (int c = 0; c < (Nx - 1) * (Nz - 1); c++) {
        eqn_p[c] = -1;
        if (g->BCp_t[c] == -1) { eqn_p[c] = inc++; incp++; }
    }
    *neq_cont = incp;
    return inc;
}

/* Node set selected by `centroid` (ParticleRoutines.c:2190-2213). */
static void node_set(const orc_grid *g, int centroid, int *nx, int *nz, const double **X,
                     const double **Z) {
    switch (centroid) {
    case 1:  *nx = g->Nx - 1; *nz = g->Nz - 1; *X = g->xc;  *Z = g->zc;  break;
    case 0:  *nx = g->Nx;     *nz = g->Nz;     *X = g->xg;  *Z = g->zg;  break;
    case -1: *nx = g->Nx;     *nz = g->Nz + 1; *X = g->xg;  *Z = g->zvx; break;
    default: *nx = g->Nx + 1; *nz = g->Nz;     *X = g->xvz; *Z = g->zg;  break;
    }
}

/* ParticleRoutines.c:2300-2310: nearest node, ties to the lower index. */
static inline int near_index(double pos, double origin, double h, int n) {
    const double distance = pos - origin;
    int i = (int)ceil(distance / h + 0.5) - 1;
    if (i < 0) i = 0;
    if (i > n - 1) i = n - 1;
    return i;
}

void orc_marker_index(const orc_grid *g, int centroid, int n, const double *x, const double *z,
                      int *ip, int *jp) {
    int nx, nz;
    const double *X, *Z;
    node_set(g, centroid, &nx, &nz, &X, &Z);
    for (int k = 0; k < n; k++) {
        ip[k] = near_index(x[k], X[0], g->dx, nx);
        jp[k] = near_index(z[k], Z[0], g->dz, nz);
    }
}

/* One weight contribution (the 'Center' block and its 8 neighbours of
 * ParticleRoutines.c:2357-2576), accumulated into per-thread buffers or with
 * atomics (mode 3, ParticleRoutines.c:2366-2371). */
static inline void p2g_add(int mode, double *Wm, double *BmWm, double *Wph, int nn, int p,
                           int kp, double w, double val, int prop) {
    if (mode == 3) {
#pragma omp atomic
        Wm[kp] += w;
#pragma omp atomic
        BmWm[kp] += val * w;
        if (prop == 1) {
#pragma omp atomic
            Wph[(size_t)p * nn + kp] += w;
        }
    } else {
        Wm[kp] += w;
        BmWm[kp] += val * w;
        if (prop == 1) Wph[(size_t)p * nn + kp] += w;
    }
}

void orc_p2mastah(const orc_grid *g, int n, const double *x, const double *z, const int *phase,
                  const double *mat_prop, int flag, int avg, int prop, int centroid, int stencil,
                  int periodic_x, const signed char *node_type, double *node_field,
                  int nb_phases, double *phase_perc, int mode, int nthreads) {
    int nx, nz;
    const double *X, *Z;
    node_set(g, centroid, &nx, &nz, &X, &Z);
    const int nn = nx * nz;
    const double dx = g->dx, dz = g->dz;
    const double dx_itp = (stencil == 9) ? 3.0 * dx / 2.0 : dx / 2.0;
    const double dz_itp = (stencil == 9) ? 3.0 * dz / 2.0 : dz / 2.0;
    const int nph = (prop == 1) ? nb_phases : 0;
    if (nthreads < 1) nthreads = 1;
    if (mode == 3) nthreads = nthreads;  /* shared buffers, atomics */
    const int nbuf = (mode == 3) ? 1 : nthreads;

    double *Wm = (double *)calloc((size_t)nbuf * nn, sizeof(double));
    double *BmWm = (double *)calloc((size_t)nbuf * nn, sizeof(double));
    double *Wph = (double *)calloc((size_t)nbuf * (nph ? nph : 1) * nn, sizeof(double));

#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
        const int tid = (mode == 3) ? 0 : omp_get_thread_num();
#else
        const int tid = 0;
#endif
        double *wm = Wm + (size_t)tid * nn;
        double *bm = BmWm + (size_t)tid * nn;
        double *wp = Wph + (size_t)tid * (nph ? nph : 1) * nn;
#pragma omp for schedule(static)
        for (int k = 0; k < n; k++) {
            if (phase[k] == -1) continue;
            const int p = phase[k];
            const int ip = near_index(x[k], X[0], dx, nx);
            const int jp = near_index(z[k], Z[0], dz, nz);
            double val = 0.0;
            if (prop == 0) {
                val = (flag == 0) ? mat_prop[p] : mat_prop[k];
                if (avg == 1) val = 1.0 / val;
                if (avg == 2) val = log(val);
            }
            double dxm = fabs(X[ip] - x[k]);
            double dzm = fabs(Z[jp] - z[k]);
            p2g_add(mode, wm, bm, wp, nn, p, ip + jp * nx,
                    (1.0 - dxm / dx_itp) * (1.0 - dzm / dz_itp), val, prop);
            if (stencil != 9) continue;
            /* 8 neighbours, in the reference order N, S, E, W, SW, NW, NE, SE.
             * Positions use X[ip] +- dx (not X[ip+-1]); x wraps when periodic. */
            const int ox[8] = {0, 0, 1, -1, -1, -1, 1, 1};
            const int oz[8] = {1, -1, 0, 0, -1, 1, 1, -1};
            for (int q = 0; q < 8; q++) {
                const int jn = jp + oz[q];
                if (jn < 0 || jn > nz - 1) continue;
                int in = ip + ox[q];
                if (in < 0 || in > nx - 1) {
                    if (!periodic_x) continue;
                    in = (in < 0) ? nx - 1 : 0;
                }
                dxm = fabs(X[ip] + ox[q] * dx - x[k]);
                dzm = fabs(Z[jp] + oz[q] * dz - z[k]);
                if (ox[q] == 0) dxm = fabs(X[ip] - x[k]);
                if (oz[q] == 0) dzm = fabs(Z[jp] - z[k]);
                p2g_add(mode, wm, bm, wp, nn, p, in + jn * nx,
                        (1.0 - dxm / dx_itp) * (1.0 - dzm / dz_itp), val, prop);
            }
        }
    }
    /* Serial-order reduction of thread buffers (ParticleRoutines.c:2582-2596). */
    double *WM = Wm, *BMWM = BmWm;
    if (nbuf > 1) {
        for (int i = 0; i < nn; i++) {
            double a = 0.0, b = 0.0;
            for (int t = 0; t < nbuf; t++) {
                a += Wm[(size_t)t * nn + i];
                b += BmWm[(size_t)t * nn + i];
            }
            Wm[i] = a;
            BmWm[i] = b;
            for (int p = 0; p < nph; p++) {
                double s = 0.0;
                for (int t = 0; t < nbuf; t++) s += Wph[((size_t)t * nph + p) * nn + i];
                Wph[(size_t)p * nn + i] = s;
            }
        }
    }
    if (prop == 1)
        for (int p = 0; p < nph; p++)
            for (int i = 0; i < nn; i++) phase_perc[(size_t)p * nn + i] = Wph[(size_t)p * nn + i];
    /* Node values (ParticleRoutines.c:2607-2631): untouched where WM<1e-30
     * or the node is air / surface-pressure tagged. */
    for (int i = 0; i < nn; i++) {
        if (fabs(WM[i]) < 1e-30 || (node_type && (node_type[i] == 30 || node_type[i] == 31)))
            continue;
        if (prop == 0) {
            double v = BMWM[i] / WM[i];
            if (avg == 1) v = 1.0 / v;
            if (avg == 2) v = exp(v);
            node_field[i] = v;
        } else {
            for (int p = 0; p < nph; p++) phase_perc[(size_t)p * nn + i] /= WM[i];
        }
    }
    /* Periodic vertex average (ParticleRoutines.c:2634-2641). */
    if (centroid == 0 && prop == 0 && periodic_x == 1)
        for (int l = 0; l < nz; l++) {
            const int c1 = l * nx + nx - 1;
            const double av = 0.5 * (node_field[c1] + node_field[l * nx]);
            node_field[c1] = av;
            node_field[l * nx] = av;
        }
    free(Wm);
    free(BmWm);
    free(Wph);
}

/* GridRoutines.c:108-212: ghost-padded centroid array, zero-flux sides,
 * quarter-sum on vertices with air/surface neighbours masked (divisor 4). */
void orc_c2v(const orc_grid *g, int per, const double *cent, double *vert) {
    const int Nx = g->Nx, Nz = g->Nz, Ncx = Nx - 1, Ncz = Nz - 1, W = Ncx + 2;
    double *t = (double *)calloc((size_t)W * (Ncz + 2), sizeof(double));
    int *f = (int *)calloc((size_t)W * (Ncz + 2), sizeof(int));
    for (int l = 0; l < Ncz; l++)
        for (int k = 0; k < Ncx; k++) {
            t[k + 1 + (l + 1) * W] = cent[k + l * Ncx];
            f[k + 1 + (l + 1) * W] = g->BCp_t[k + l * Ncx];
        }
#define CP(dst, src) do { t[dst] = t[src]; f[dst] = f[src]; } while (0)
    for (int k = 1; k < Ncx + 1; k++) CP(k, k + W);
    for (int k = 1; k < Ncx + 1; k++) CP(k + (Ncz + 1) * W, k + Ncz * W);
    for (int l = 1; l < Ncz + 1; l++) CP(l * W, (per ? W - 2 : 1) + l * W);
    for (int l = 1; l < Ncz + 1; l++) CP(Ncx + 1 + l * W, (per ? 1 : Ncx) + l * W);
    CP(0, per ? W : 1 + W);
    CP(Ncx + 1, per ? Ncx + 1 + W : Ncx + W);
    CP((Ncz + 1) * W, per ? Ncz * W : 1 + Ncz * W);
    CP(Ncx + 1 + (Ncz + 1) * W, per ? Ncx + 1 + Ncz * W : Ncx + Ncz * W);
#undef CP
    for (int k = 0; k < Nx; k++)
        for (int l = 0; l < Nz; l++) {
            const int c1 = k + l * Nx, c0 = k + (l + 1) * W + 1;
            vert[c1] = 0.0;
            if (g->BCg_t[c1] == 30) continue;
            const int sw = c0 - 1 - W, se = c0 - W, nw = c0 - 1, ne = c0;
            const int iSW = !(f[sw] == 30 || f[sw] == 31), iSE = !(f[se] == 30 || f[se] == 31);
            const int iNW = !(f[nw] == 30 || f[nw] == 31), iNE = !(f[ne] == 30 || f[ne] == 31);
            vert[c1] = 0.25 * (iSW * t[sw] + iSE * t[se] + iNW * t[nw] + iNE * t[ne]);
        }
    free(t);
    free(f);
}

/* StokesRoutines.c:415-500: ghost velocities from the tags. */
void orc_apply_bc(const orc_grid *g, double *u, double *v) {
    const int nx = g->Nx, nz = g->Nz, nzvx = nz + 1, nxvz = nx + 1;
    for (int i = 0; i < nx; i++) {
        if (g->BCu_t[i] == 13) u[i] = u[i + nx];
        const int iN = i + (nzvx - 1) * nx;
        if (g->BCu_t[iN] == 13) u[iN] = u[iN - nx];
    }
    for (int i = 0; i < nx; i++) {
        if (g->BCu_t[i] == 11) u[i] = g->BCu_v[i] - u[i + nx];
        const int iN = i + (nzvx - 1) * nx;
        if (g->BCu_t[iN] == 11) u[iN] = g->BCu_v[iN] - u[iN - nx];
    }
    for (int j = 0; j < nz; j++) {
        const int w = j * nxvz, e = j * nxvz + nxvz - 1;
        if (g->BCv_t[w] == 13) v[w] = v[w + 1];
        if (g->BCv_t[e] == 13) v[e] = v[e - 1];
    }
    for (int j = 0; j < nz; j++) {
        const int w = j * nxvz, e = j * nxvz + nxvz - 1;
        if (g->BCv_t[w] == 11) v[w] = g->BCv_v[w] - v[w + 1];
        if (g->BCv_t[e] == 11) v[e] = g->BCv_v[e] - v[e - 1];
    }
    for (int j = 0; j < nz; j++) {
        const int w = j * nxvz, e = j * nxvz + nxvz - 1;
        if (g->BCv_t[w] == -12) v[w] = v[e - 1];
        if (g->BCv_t[e] == -12) v[e] = v[w + 1];
    }
    for (int j = 0; j < nzvx; j++)
        if (g->BCu_t[j * nx + nx - 1] == -12) u[j * nx + nx - 1] = u[j * nx];
}

/* RheologyDensity.c:2001-2097 (out_of_plane = 0). */
void orc_strain_rates(const orc_grid *g, int compressible, const double *u, const double *v,
                      double *div, double *exxd, double *ezzd, double *exz, double *wxz,
                      double *exxd_s, double *ezzd_s, double *div_s) {
    const int Nx = g->Nx, Nz = g->Nz, Ncx = Nx - 1, Ncz = Nz - 1;
    const double dx = g->dx, dz = g->dz;
    for (int l = 0; l < Ncz; l++)
        for (int k = 0; k < Ncx; k++) {
            const int c0 = k + l * Ncx, c1 = k + l * Nx, c2 = k + l * (Nx + 1);
            div[c0] = exxd[c0] = ezzd[c0] = 0.0;
            if (g->BCp_t[c0] == 30 || g->BCp_t[c0] == 31) continue;
            const double dvxdx = (u[c1 + 1 + Nx] - u[c1 + Nx]) / dx;
            const double dvzdz = (v[c2 + Nx + 1 + 1] - v[c2 + 1]) / dz;
            div[c0] = dvxdx + dvzdz;
            exxd[c0] = dvxdx - compressible * 1.0 / 3.0 * div[c0];
            ezzd[c0] = dvzdz - compressible * 1.0 / 3.0 * div[c0];
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx; k++) {
            const int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
            exz[c1] = 0.0;
            if (g->BCg_t[c1] == 30) continue;
            const double dvxdz = (u[c1 + Nx] - u[c1]) / dz;
            const double dvzdx = (v[c2 + 1] - v[c2]) / dx;
            exz[c1] = 0.5 * (dvxdz + dvzdx);
            if (wxz) wxz[c1] = 0.5 * (dvzdx - dvxdz);
        }
    if (exxd_s) orc_c2v(g, 0, exxd, exxd_s);
    if (ezzd_s) orc_c2v(g, 0, ezzd, ezzd_s);
    if (div_s) orc_c2v(g, 0, div, div_s);
}

/* StokesRoutines.c:983-1100 (polar = 0, no surface processes): the body
 * force at a velocity node is g times the mean density of its two
 * neighbours (vertices for Vx, cells for Vz); the stored value is +g*rho. */
void orc_eval_rhs(const orc_grid *g, double gx, double gz, const double *rho_n,
                  const double *rho_s, double *roger_x, double *roger_z) {
    const int Nx = g->Nx, Nz = g->Nz, NzVx = Nz + 1, NxVz = Nx + 1, Ncx = Nx - 1;
    for (int l = 0; l < NzVx; l++)
        for (int k = 0; k < Nx; k++) {
            const int c = k + l * Nx;
            double r = 0.0;
            if (l > 0 && l < NzVx - 1) {
                const signed char t = g->BCu_t[c];
                if (t == -1 || t == 2 || t == -2)
                    r = -gx * (0.5 * (rho_s[c - Nx] + rho_s[c]));
            }
            roger_x[c] = -r;
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < NxVz; k++) {
            const int c = k + l * NxVz, c2 = k - 1 + (l - 1) * Ncx;
            double r = 0.0;
            if (k > 0 && k < NxVz - 1 && g->BCv_t[c] == -1)
                r = -gz * (0.5 * (rho_n[c2] + rho_n[c2 + Ncx]));
            roger_z[c] = -r;
        }
}

/* RheologyDensity.c:2066-2085: exz_n = 1/4 of the non-air vertex values
 * around each non-air cell. */
void orc_exz_n(const orc_grid *g, const double *exz, double *exz_n) {
    const int Nx = g->Nx, Ncx = Nx - 1, Ncz = g->Nz - 1;
    for (int l = 0; l < Ncz; l++)
        for (int k = 0; k < Ncx; k++) {
            const int c0 = k + l * Ncx, c1 = k + l * Nx;
            double s = 0.0;
            if (g->BCp_t[c0] != 30 && g->BCp_t[c0] != 31) {
                if (g->BCg_t[c1] != 30) s += 0.25 * exz[c1];
                if (g->BCg_t[c1 + 1] != 30) s += 0.25 * exz[c1 + 1];
                if (g->BCg_t[c1 + Nx] != 30) s += 0.25 * exz[c1 + Nx];
                if (g->BCg_t[c1 + Nx + 1] != 30) s += 0.25 * exz[c1 + Nx + 1];
            }
            exz_n[c0] = s;
        }
}
