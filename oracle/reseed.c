/*
 * Oracle: marker re-seeding of CountPartCell_OLD (ParticleRoutines.c:3224-
 * 4050, reseed_mode = 0, the default), single-threaded as the reference runs
 * it (nthreads = 1), for a box without free surface and without periodic
 * x.  TEST INFRASTRUCTURE ONLY.
 *
 * The box is covered by an extended grid of one extra coarse cell on every
 * side, split into fine cells of dx/2 x dz/2 (2 (Ncx+2) x 2 (Ncz+2)).  Each
 * active marker is counted in the fine cell of its nearest fine-cell centre
 * (:3626-3635).  Every interior fine cell whose coarse cell is not air
 * (BCt 30; the air tag is shared with BCp) and that holds fewer than
 * floor(min_part_cell / 4) markers (:3767) receives one new marker at its
 * centre (AddPartCell2, :1048-1094), which copies the phase (and the other
 * carried fields) of the closest existing marker among the 5 x 5 fine cells
 * around it, taken in the reference's list order with strict '<'
 * (FindClosestPhase2, :963-1006; list :3772-3951).  New markers take the
 * inactive slots (phase -1) in increasing order first, then are appended
 * (:3972-4021).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "mdoracle.h"

static int fine_index(double pos, double origin, double h, int n) {
    const double distance = pos - origin;
    int i = (int)ceil((distance / h * 2.0) + 0.5) - 1;
    if (i < 0) i = 0;
    if (i > n - 1) i = n - 1;
    return i;
}

/* Returns the number of new markers (>= 0), -1 when a deficient fine cell
 * has no marker in its 5 x 5 neighbourhood (the reference exits, :3812),
 * -2 when the new markers do not fit in capacity (:4011).  On success x, z
 * and phase hold the new markers at their slots, new_slot / new_src (room
 * for one entry per fine cell) list each new marker's slot and the marker it
 * copied, and *n_out is the new marker count including appended slots. */
int orc_reseed(const orc_grid *g, double xmin, double xmax, int min_part_cell, int n, int capacity, double *x,
               double *z, int *phase, int *new_slot, int *new_src, int *n_out) {
    const int Ncx = g->Nx - 1, Ncz = g->Nz - 1, ncx_e = Ncx + 2, ncz_e = Ncz + 2;
    const int nfx = 2 * ncx_e, nfz = 2 * ncz_e, nf = nfx * nfz;
    const double dx = g->dx, dz = g->dz;
    /* extended coordinates (:3280-3289, :3368-3375): running sums */
    const double xg_e0 = g->xg[0] - dx, xg_e1 = xg_e0 + dx;
    const double zg_e0 = g->zg[0] - dz, zg_e1 = zg_e0 + dz;
    const double xc_e0 = 0.5 * (xg_e0 + xg_e1), zc_e0 = 0.5 * (zg_e0 + zg_e1);
    int *mpc = (int *)calloc((size_t)nf, sizeof(int));
    int *off = (int *)calloc((size_t)nf + 1, sizeof(int));
    int *cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *lst = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    int *fill = (int *)calloc((size_t)nf, sizeof(int));
    int *ind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    double *nx_ = (double *)malloc(sizeof(double) * (size_t)nf), *nz_ = (double *)malloc(sizeof(double) * (size_t)nf);
    int nnew = 0, rc = 0;
    for (int k = 0; k < n; k++) {
        cell[k] = -1;
        if (phase[k] == -1) continue;
        const int ic = fine_index(x[k], xc_e0 - dx / 4.0, dx, nfx);
        const int jc = fine_index(z[k], zc_e0 - dz / 4.0, dz, nfz);
        cell[k] = ic + jc * nfx;
        mpc[cell[k]]++;
    }
    /* per-cell marker lists in increasing marker index (:3866-3881) */
    for (int q = 0; q < nf; q++) off[q + 1] = off[q] + mpc[q];
    for (int k = 0; k < n; k++)
        if (cell[k] >= 0) lst[off[cell[k]] + fill[cell[k]]++] = k;
    /* window order of the reference's ind_list (:3783-3808, :3820-3951) */
    const int nxl = nfx;
    const int win[25] = {0, -1, +1, -nxl, +nxl, -1 - nxl, +1 - nxl, -1 + nxl, +1 + nxl,
                         -2, +2, -2 * nxl, +2 * nxl, -2 - 2 * nxl, +2 - 2 * nxl, -2 + 2 * nxl, +2 + 2 * nxl,
                         -1 - 2 * nxl, +1 - 2 * nxl, -1 + 2 * nxl, +1 + 2 * nxl, -2 - nxl, -2 + nxl, +2 - nxl,
                         +2 + nxl};
    const int thr = (int)(ceil(min_part_cell) / 4.0);
    for (int k = 2; k < nfx - 2 && rc == 0; k++) {
        for (int l = 2; l < nfz - 2; l++) {
            const int kd = k + l * nfx;
            double distance = (xg_e0 + dx / 4.0 + (k)*dx / 2.0) - g->xc[0];
            int ic = (int)ceil((distance / dx) + 0.5) - 1;
            if (ic < 0) ic = 0;
            if (ic > Ncx - 1) ic = Ncx - 1;
            distance = (zg_e0 + dz / 4.0 + (l)*dz / 2.0) - g->zc[0];
            int jc = (int)ceil((distance / dz) + 0.5) - 1;
            if (jc < 0) jc = 0;
            if (jc > Ncz - 1) jc = Ncz - 1;
            if (g->BCp_t[ic + jc * Ncx] == 30) continue;
            if (mpc[kd] >= thr) continue;
            int nb = 0;
            for (int w = 0; w < 25; w++)
                for (int m = off[kd + win[w]]; m < off[kd + win[w] + 1]; m++) ind[nb++] = lst[m];
            if (nb == 0) { rc = -1; break; }
            const double new_x = g->xg[0] + (k - 2) * dx / 2.0 + dx / 4.0;
            const double new_z = g->zg[0] + (l - 2) * dz / 2.0 + dz / 4.0;
            double min_dst = xmax - xmin;
            int min_index = ind[0];
            for (int i = 0; i < nb; i++) {
                const double ddx = new_x - x[ind[i]], ddz = new_z - z[ind[i]];
                const double dst = sqrt(ddx * ddx + ddz * ddz);
                if (dst < min_dst) { min_dst = dst; min_index = ind[i]; }
            }
            nx_[nnew] = new_x;
            nz_[nnew] = new_z;
            new_src[nnew] = min_index;
            nnew++;
        }
    }
    if (rc == 0) {
        /* slots: inactive markers in increasing index, then appended */
        int reuse = 0, appended = 0;
        for (int i = 0, k = 0; i < nnew; i++) {
            int slot;
            while (k < n && phase[k] != -1) k++;
            if (k < n) { slot = k; k++; reuse++; }
            else slot = n + appended++;
            if (slot >= capacity) { rc = -2; break; }
            new_slot[i] = slot;
        }
        if (rc == 0) {
            for (int i = 0; i < nnew; i++) {
                const int s = new_slot[i];
                x[s] = nx_[i];
                z[s] = nz_[i];
                phase[s] = phase[new_src[i]];
            }
            *n_out = n + appended;
            rc = nnew;
        }
    }
    free(mpc); free(off); free(cell); free(lst); free(fill); free(ind); free(nx_); free(nz_);
    return rc;
}
