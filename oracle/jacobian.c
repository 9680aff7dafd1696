/*
 * Oracle: assembled Newton Jacobian of the Stokes momentum rows.
 * TEST INFRASTRUCTURE ONLY.
 *
 * Restates Xjacobian_InnerNodesDecoupled3 (StokesAssemblyDecoupled.c:219-556),
 * Zjacobian_InnerNodesDecoupled3 (:732-1085) and the row loops of
 * BuildJacobianOperatorDecoupled (:1786-2304) for inner nodes without
 * periodicity, stress boundaries or density stabilisation.  The continuity
 * block is the Picard C of orc_assemble_picard.
 *
 * The reference's coefficients are sympy expressions in D11..D34, the
 * node flags and comp*out_of_plane (BuildJacobianOperatorDecoupled fixes
 * comp = 1).  They are regrouped here into the responses of the tangent
 * stresses to the deviatoric normal strain rates (exx_d = exx - co (exx +
 * ezz)/3, co = comp*out_of_plane, cm = co - 3):
 *   on centres   a = (cm D11 + co D12)/3,  b = (co D11 + cm D12)/3   (x rows)
 *                a2 = (co D21 + cm D22)/3, b2 = (cm D21 + co D22)/3  (z rows)
 *   on vertices  s = (cm D31 + co D32)/3,  t = (co D31 + cm D32)/3
 * (so -d sxx/d exx = a, -d sxx/d ezz = b, ...).  Each coefficient below is
 * the reference's, rewritten in these terms.
 */
#include <stdlib.h>
#include <string.h>

#include "mdoracle.h"

static int folds(int t) { return t == 0 || t == 31 || t == 11 || t == 13; }
static int free_u(int t) { return t == -1 || t == 2 || t == -2; }
static double flag(int cond) { return cond ? 1.0 : 0.0; }

typedef struct {
    int *p, *j;
    double *x;
    int n;
} csr_b;

/* AddCoeff3 (StokesAssemblyDecoupled.c:46-56) for the matrix part: Dirichlet
 * type columns go to the right-hand side, the others become entries. */
static void put(csr_b *m, int col_type, int col_eqn, double coef) {
    if (folds(col_type)) return;
    m->j[m->n] = col_eqn;
    m->x[m->n] = coef;
    m->n++;
}

void orc_assemble_jacobian(const orc_grid *g, const int *eqn_u, const int *eqn_v, const int *eqn_p,
                           int neq_mom, int neq_cont, const double *Dn, const double *Ds, double co,
                           int *Ap, int *Aj, double *Ax, int *Bp, int *Bj, double *Bx) {
    (void)neq_cont;
    const int nx = g->Nx, nz = g->Nz, nzvx = nz + 1, nxvz = nx + 1, ncx = nx - 1;
    const int nc = (nx - 1) * (nz - 1), nv = nx * nz;
    const double dx = g->dx, dz = g->dz, celvol = dx * dz, cm = co - 3.0;
    const signed char *Ut = g->BCu_t, *Vt = g->BCv_t, *Pt = g->BCp_t, *Gt = g->BCg_t;
    const double *D11 = Dn, *D12 = Dn + nc, *D13 = Dn + 2 * nc, *D14 = Dn + 3 * nc;
    const double *D21 = Dn + 4 * nc, *D22 = Dn + 5 * nc, *D23 = Dn + 6 * nc, *D24 = Dn + 7 * nc;
    const double *D31 = Ds, *D32 = Ds + nv, *D33 = Ds + 2 * nv, *D34 = Ds + 3 * nv;
    csr_b A = {Ap, Aj, Ax, 0}, B = {Bp, Bj, Bx, 0};
    int row = 0;
    const double dx2 = dx * dx, dz2 = dz * dz, dxz = dx * dz;

    /* ------------------------------------------------------- x momentum */
    for (int l = 0; l < nzvx; l++)
        for (int k = 0; k < nx; k++) {
            const int c1 = k + l * nx;
            if (!free_u(Ut[c1])) continue;
            Ap[row] = A.n; Bp[row] = B.n; row++;
            if (!(l > 0 && l < nzvx - 1 && k > 0 && k < nx - 1)) continue;
            const int c2 = k - 1 + (l - 1) * ncx, c3 = k + l * nxvz;
            const int iVxS = c1 - nx, iVxN = c1 + nx, iVxW = c1 - 1, iVxE = c1 + 1;
            const int iVxSW = iVxS - 1, iVxSE = iVxS + 1, iVxNW = iVxN - 1, iVxNE = iVxN + 1;
            const int iVzSW = c3 - nxvz, iVzSE = c3 - nxvz + 1, iVzNW = c3, iVzNE = c3 + 1;
            const int iVzSWW = iVzSW - 1, iVzSEE = iVzSE + 1, iVzNWW = iVzNW - 1, iVzNEE = iVzNE + 1;
            const int iVzNNW = iVzNW + nxvz, iVzSSW = iVzSW - nxvz, iVzNNE = iVzNE + nxvz, iVzSSE = iVzSE - nxvz;
            const int iPrW = c2, iPrE = c2 + 1, iPrSW = iPrW - ncx, iPrSE = iPrE - ncx, iPrNW = iPrW + ncx,
                      iPrNE = iPrE + ncx;
            const int ixyN = c1, ixyS = c1 - nx;
            int ixySW = ixyS - 1, ixySE = ixyS + 1, ixyNW = ixyN - 1, ixyNE = ixyN + 1;
            if (k == 0) { ixySW = ixyS; ixyNW = ixyN; }
            if (k == nx - 1) { ixySE = ixyS; ixyNE = ixyN; }
            /* flags (:318-374) */
            const double inW = flag(Pt[iPrW] == -1), inE = flag(Pt[iPrE] == -1);
            const double inS = flag(Ut[iVxS] != 30 && Ut[iVxS] != 13), inN = flag(Ut[iVxN] != 30 && Ut[iVxN] != 13);
            const double inSv = flag(Gt[ixyS] != 30), inNv = flag(Gt[ixyN] != 30);
            double inSWc = 0, inSEc = 0, inNWc = 0, inNEc = 0, inSWv = 0, inSEv = 0, inNWv = 0, inNEv = 0;
            if (l > 1) { inSWc = flag(Pt[iPrSW] == -1); inSEc = flag(Pt[iPrSE] == -1); }
            if (l < nzvx - 2) { inNWc = flag(Pt[iPrNW] == -1); inNEc = flag(Pt[iPrNE] == -1); }
            if (k > 0 || (Vt[iVzSW] == -1 || Vt[iVzSW] == 0)) { inSWv = flag(Gt[ixySW] != 30); inNWv = flag(Gt[ixyNW] != 30); }
            if (k < nx - 1 || (Vt[iVzSE] == -1 || Vt[iVzSE] == 0)) { inSEv = flag(Gt[ixySE] != 30); inNEv = flag(Gt[ixyNE] != 30); }
            const double inSWW = flag(Vt[iVzSWW] == -1), inNWW = flag(Vt[iVzNWW] == -1);
            const double inSEE = flag(Vt[iVzSEE] == -1), inNEE = flag(Vt[iVzNEE] == -1);
            /* vertex averaging weights: 1 / number of contributing neighbours */
            double wS = inW + inE + inSWc + inSEc, wN = inW + inE + inNWc + inNEc;
            if (wS > 1.0) wS = 1.0 / wS;
            if (wN > 1.0) wN = 1.0 / wN;
            /* stress responses */
            const double aW = (cm * D11[iPrW] + co * D12[iPrW]) / 3.0, aE = (cm * D11[iPrE] + co * D12[iPrE]) / 3.0;
            const double bW = (co * D11[iPrW] + cm * D12[iPrW]) / 3.0, bE = (co * D11[iPrE] + cm * D12[iPrE]) / 3.0;
            const double sS = (cm * D31[ixyS] + co * D32[ixyS]) / 3.0, sN = (cm * D31[ixyN] + co * D32[ixyN]) / 3.0;
            const double tS = (co * D31[ixyS] + cm * D32[ixyS]) / 3.0, tN = (co * D31[ixyN] + cm * D32[ixyN]) / 3.0;
            const double D13W = D13[iPrW], D13E = D13[iPrE], D14W = D14[iPrW], D14E = D14[iPrE];
            const double D33S = D33[ixyS], D33N = D33[ixyN], D34S = D34[ixyS], D34N = D34[ixyN];
            const double qS = inS * inS * inSv, qN = inN * inN * inNv;   /* active shear-stress vertices */
            const double shear = wS * sS - wN * sN;
            double uW = inW * aW / dx2 + inW * shear / dxz - 0.25 * D13W * (inN * inNWv - inS * inSWv) / dxz;
            double uC = (D33N * qN + D33S * qS) / dz2 + (inE - inW) * shear / dxz -
                        0.25 * (D13W - D13E) * (qN - qS) / dxz - (inE * aE + inW * aW) / dx2;
            double uE = inE * aE / dx2 - inE * shear / dxz + 0.25 * D13E * (inN * inNEv - inS * inSEv) / dxz;
            const double uS = -D33S * qS / dz2 + 0.25 * qS * (D13E - D13W) / dxz + wS * (inSEc - inSWc) * sS / dxz;
            const double uN = -D33N * qN / dz2 + 0.25 * qN * (D13W - D13E) / dxz - wN * (inNEc - inNWc) * sN / dxz;
            const double uSW = (-0.25 * D13W * inS * inSWv + inSWc * wS * sS) / dxz;
            const double uSE = (0.25 * D13E * inS * inSEv - inSEc * wS * sS) / dxz;
            const double uNW = (0.25 * D13W * inN * inNWv - inNWc * wN * sN) / dxz;
            const double uNE = (-0.25 * D13E * inN * inNEv + inNEc * wN * sN) / dxz;
            double vSW = -D33S * inS * inSv / dxz - (inW * wN * tN + (inSWc - inW) * wS * tS) / dz2 + inW * bW / dxz +
                         0.25 * (D13E * inS * inSv - D13W * (inS * inSv - inSWW * inSWv)) / dx2;
            double vSE = D33S * inS * inSv / dxz + (-inE * wN * tN + (inE - inSEc) * wS * tS) / dz2 - inE * bE / dxz -
                         0.25 * (D13E * (inS * inSv - inSEE * inSEv) - D13W * inS * inSv) / dx2;
            double vNW = D33N * inN * inNv / dxz - ((inNWc - inW) * wN * tN + inW * wS * tS) / dz2 - inW * bW / dxz -
                         0.25 * (-D13E * inN * inNv + D13W * (inN * inNv - inNWW * inNWv)) / dx2;
            double vNE = -D33N * inN * inNv / dxz - (-(inE - inNEc) * wN * tN + inE * wS * tS) / dz2 + inE * bE / dxz +
                         0.25 * (-D13E * (inN * inNv - inNEE * inNEv) + D13W * inN * inNv) / dx2;
            const double vSWW = -0.25 * D13W * inSWW * inSWv / dx2, vSEE = -0.25 * D13E * inSEE * inSEv / dx2;
            const double vNWW = -0.25 * D13W * inNWW * inNWv / dx2, vNEE = -0.25 * D13E * inNEE * inNEv / dx2;
            const double vSSW = inSWc * wS * tS / dz2, vSSE = inSEc * wS * tS / dz2;
            const double vNNW = inNWc * wN * tN / dz2, vNNE = inNEc * wN * tN / dz2;
            const double pshear = (D34S * inS * wS - D34N * inN * wN) / dz;
            const double pW = -inW / dx + inW * (D14W / dx + pshear), pE = inE / dx + inE * (-D14E / dx + pshear);
            const double pSW = D34S * inS * inSWc * wS / dz, pSE = D34S * inS * inSEc * wS / dz;
            const double pNW = -D34N * inN * inNWc * wN / dz, pNE = -D34N * inN * inNEc * wN / dz;
            /* no-slip ghosts (:420-429): ghost = -interior folds into the interior coefficient */
            if (Ut[iVxS] == 11) uC -= uS;
            if (Ut[iVxN] == 11) uC -= uN;
            if (Ut[iVxSW] == 11) uW -= uSW;
            if (Ut[iVxSE] == 11) uE -= uSE;
            if (Ut[iVxNW] == 11) uW -= uNW;
            if (Ut[iVxNE] == 11) uE -= uNE;
            if (Vt[iVzNWW] == 11) vNW -= vNWW;
            if (Vt[iVzNEE] == 11) vNE -= vNEE;
            if (Vt[iVzSWW] == 11) vSW -= vSWW;
            if (Vt[iVzSEE] == 11) vSE -= vSEE;
            /* entries in the reference's order (:431-527) */
            if (Ut[iVxSW] != 30) put(&A, Ut[iVxSW], eqn_u[iVxSW], uSW * celvol);
            if (Ut[iVxS] != 30) put(&A, Ut[iVxS], eqn_u[iVxS], uS * celvol);
            if (Ut[iVxSE] != 30) put(&A, Ut[iVxSE], eqn_u[iVxSE], uSE * celvol);
            if (Ut[iVxW] != 30) put(&A, Ut[iVxW], eqn_u[iVxW], uW * celvol);
            if (Ut[c1] != 30) put(&A, Ut[c1], eqn_u[c1], uC * celvol);
            if (Ut[iVxE] != 30) put(&A, Ut[iVxE], eqn_u[iVxE], uE * celvol);
            if (Ut[iVxNW] != 30) put(&A, Ut[iVxNW], eqn_u[iVxNW], uNW * celvol);
            if (Ut[iVxN] != 30) put(&A, Ut[iVxN], eqn_u[iVxN], uN * celvol);
            if (Ut[iVxNE] != 30) put(&A, Ut[iVxNE], eqn_u[iVxNE], uNE * celvol);
            if (l > 1 && Vt[iVzSSW] != 30) put(&A, Vt[iVzSSW], eqn_v[iVzSSW], vSSW * celvol);
            if (l > 1 && Vt[iVzSSE] != 30) put(&A, Vt[iVzSSE], eqn_v[iVzSSE], vSSE * celvol);
            if ((k > 1 || Vt[iVzSWW] == -1) && Vt[iVzSWW] != 30) put(&A, Vt[iVzSWW], eqn_v[iVzSWW], vSWW * celvol);
            if (Vt[iVzSW] != 30) put(&A, Vt[iVzSW], eqn_v[iVzSW], vSW * celvol);
            if (Vt[iVzSE] != 30) put(&A, Vt[iVzSE], eqn_v[iVzSE], vSE * celvol);
            if ((k < nx - 2 || Vt[iVzSEE] == -1) && Vt[iVzSEE] != 30) put(&A, Vt[iVzSEE], eqn_v[iVzSEE], vSEE * celvol);
            if ((k > 1 || Vt[iVzNWW] == -1) && Vt[iVzNWW] != 30) put(&A, Vt[iVzNWW], eqn_v[iVzNWW], vNWW * celvol);
            if (Vt[iVzNW] != 30) put(&A, Vt[iVzNW], eqn_v[iVzNW], vNW * celvol);
            if (Vt[iVzNE] != 30) put(&A, Vt[iVzNE], eqn_v[iVzNE], vNE * celvol);
            if ((k < nx - 2 || Vt[iVzNEE] == -1) && Vt[iVzNEE] != 30) put(&A, Vt[iVzNEE], eqn_v[iVzNEE], vNEE * celvol);
            if (l < nz - 1 && Vt[iVzNNW] != 30) put(&A, Vt[iVzNNW], eqn_v[iVzNNW], vNNW * celvol);
            if (l < nz - 1 && Vt[iVzNNE] != 30) put(&A, Vt[iVzNNE], eqn_v[iVzNNE], vNNE * celvol);
            if (l > 1) {
                if (Pt[iPrSW] != 30) put(&B, Pt[iPrSW], eqn_p[iPrSW] - neq_mom, pSW * celvol);
                if (Pt[iPrSE] != 30) put(&B, Pt[iPrSE], eqn_p[iPrSE] - neq_mom, pSE * celvol);
            }
            if (Pt[iPrW] != 30) put(&B, Pt[iPrW], eqn_p[iPrW] - neq_mom, pW * celvol);
            if (Pt[iPrE] != 30) put(&B, Pt[iPrE], eqn_p[iPrE] - neq_mom, pE * celvol);
            if (l < nz - 1) {
                if (Pt[iPrNW] != 30) put(&B, Pt[iPrNW], eqn_p[iPrNW] - neq_mom, pNW * celvol);
                if (Pt[iPrNE] != 30) put(&B, Pt[iPrNE], eqn_p[iPrNE] - neq_mom, pNE * celvol);
            }
        }

    /* ------------------------------------------------------- z momentum */
    for (int l = 0; l < nz; l++)
        for (int k = 0; k < nxvz; k++) {
            const int c3 = k + l * nxvz;
            if (!(Vt[c3] == -1 || Vt[c3] == 2)) continue;
            Ap[row] = A.n; Bp[row] = B.n; row++;
            if (!(k > 0 && k < nxvz - 1 && l > 0 && l < nz - 1)) continue;
            const int c1 = k + l * nx, c2 = k - 1 + (l - 1) * ncx;
            const int iVzW = c3 - 1, iVzE = c3 + 1, iVzS = c3 - nxvz, iVzN = c3 + nxvz;
            const int iVzSW = iVzS - 1, iVzSE = iVzS + 1, iVzNW = iVzN - 1, iVzNE = iVzN + 1;
            const int iVxNW = c1 + nx - 1, iVxNE = c1 + nx, iVxSW = c1 - 1, iVxSE = c1;
            const int iVxNWW = iVxNW - 1, iVxNEE = iVxNE + 1, iVxSWW = iVxSW - 1, iVxSEE = iVxSE + 1;
            const int iVxNNW = iVxNW + nx, iVxNNE = iVxNE + nx, iVxSSW = iVxSW - nx, iVxSSE = iVxSE - nx;
            const int iPrS = c2, iPrSW = iPrS - 1, iPrSE = iPrS + 1, iPrN = c2 + ncx, iPrNW = iPrN - 1,
                      iPrNE = iPrN + 1;
            const int ixyW = c1 - 1, ixySW = ixyW - nx, ixyNW = ixyW + nx, ixyE = c1, ixySE = ixyE - nx,
                      ixyNE = ixyE + nx;
            /* flags (:833-876) */
            const double inS = flag(Pt[iPrS] == -1), inN = flag(Pt[iPrN] == -1);
            const double inW = flag(Vt[iVzW] != 30 && Vt[iVzW] != 13), inE = flag(Vt[iVzE] != 30 && Vt[iVzE] != 13);
            const double inWv = flag(Gt[ixyW] != 30), inEv = flag(Gt[ixyE] != 30);
            double inSWc = 0, inSEc = 0, inNWc = 0, inNEc = 0, inSWv = 0, inSEv = 0, inNWv = 0, inNEv = 0;
            if (k > 1 || (k == 1 && Vt[iVzW] == -1)) { inSWc = flag(Pt[iPrSW] == -1); inNWc = flag(Pt[iPrNW] == -1); }
            if (k < nxvz - 2 || (k == nxvz - 2 && Vt[iVzE] == -1)) { inSEc = flag(Pt[iPrSE] == -1); inNEc = flag(Pt[iPrNE] == -1); }
            if (l > 0) { inSWv = flag(Gt[ixySW] != 30); inSEv = flag(Gt[ixySE] != 30); }
            if (l < nz - 1) { inNWv = flag(Gt[ixyNW] != 30); inNEv = flag(Gt[ixyNE] != 30); }
            const double inSSW = flag(Ut[iVxSSW] == -1), inSSE = flag(Ut[iVxSSE] == -1);
            const double inNNW = flag(Ut[iVxNNW] == -1), inNNE = flag(Ut[iVxNNE] == -1);
            double wE = inN + inS + inNEc + inSEc, wW = inN + inS + inNWc + inSWc;
            if (wW > 1.0) wW = 1.0 / wW;
            if (wE > 1.0) wE = 1.0 / wE;
            const double a2S = (co * D21[iPrS] + cm * D22[iPrS]) / 3.0, a2N = (co * D21[iPrN] + cm * D22[iPrN]) / 3.0;
            const double b2S = (cm * D21[iPrS] + co * D22[iPrS]) / 3.0, b2N = (cm * D21[iPrN] + co * D22[iPrN]) / 3.0;
            const double sW = (cm * D31[ixyW] + co * D32[ixyW]) / 3.0, sE = (cm * D31[ixyE] + co * D32[ixyE]) / 3.0;
            const double tW = (co * D31[ixyW] + cm * D32[ixyW]) / 3.0, tE = (co * D31[ixyE] + cm * D32[ixyE]) / 3.0;
            const double D23S = D23[iPrS], D23N = D23[iPrN], D24S = D24[iPrS], D24N = D24[iPrN];
            const double D33W = D33[ixyW], D33E = D33[ixyE], D34W = D34[ixyW], D34E = D34[ixyE];
            const double qW = inW * inW * inWv, qE = inE * inE * inEv;
            const double shear = wW * tW - wE * tE;
            double vW = -D33W * qW / dx2 + 0.25 * qW * (D23N * inN - D23S * inS) / dxz + wW * (inNWc - inSWc) * tW / dxz;
            double vC = -(inN * a2N + inS * a2S) / dz2 - 0.25 * (D23S * inS - D23N * inN) * (qE - qW) / dxz +
                        (inN - inS) * shear / dxz + (D33E * qE + D33W * qW) / dx2;
            double vE = -D33E * qE / dx2 + 0.25 * qE * (D23S * inS - D23N * inN) / dxz - wE * (inNEc - inSEc) * tE / dxz;
            const double vS = inS * (-0.25 * D23S * (inE * inSEv - inSWv * inW) / dxz + a2S / dz2 + shear / dxz);
            const double vN = inN * (0.25 * D23N * (inE * inNEv - inNWv * inW) / dxz + a2N / dz2 - shear / dxz);
            double uSW = 0.25 * (D23N * inN * inW * inWv + D23S * inS * (inSSW * inSWv - inW * inWv)) / dz2 +
                         inS * b2S / dxz - D33W * inW * inWv / dxz - (inS * wE * sE - (inS - inSWc) * wW * sW) / dx2;
            double uSE = -0.25 * (-D23N * inE * inEv * inN + D23S * inS * (inE * inEv - inSEv * inSSE)) / dz2 -
                         inS * b2S / dxz + D33E * inE * inEv / dxz + ((inS - inSEc) * wE * sE - inS * wW * sW) / dx2;
            double uNW = 0.25 * (D23N * inN * (inNNW * inNWv - inW * inWv) + D23S * inS * inW * inWv) / dz2 -
                         inN * b2N / dxz + D33W * inW * inWv / dxz + (-inN * wE * sE + (inN - inNWc) * wW * sW) / dx2;
            double uNE = 0.25 * (-D23N * inN * (inE * inEv - inNEv * inNNE) + D23S * inE * inEv * inS) / dz2 +
                         inN * b2N / dxz - D33E * inE * inEv / dxz - (-(inN - inNEc) * wE * sE + inN * wW * sW) / dx2;
            const double vSW = (-0.25 * D23S * inS * inSWv * inW + inSWc * wW * tW) / dxz;
            const double vSE = (0.25 * D23S * inE * inS * inSEv - inSEc * wE * tE) / dxz;
            const double vNW = (0.25 * D23N * inN * inNWv * inW - inNWc * wW * tW) / dxz;
            const double vNE = (-0.25 * D23N * inE * inN * inNEv + inNEc * wE * tE) / dxz;
            const double uSWW = inSWc * wW * sW / dx2, uSEE = inSEc * wE * sE / dx2;
            const double uNWW = inNWc * wW * sW / dx2, uNEE = inNEc * wE * sE / dx2;
            const double uSSW = -0.25 * D23S * inS * inSSW * inSWv / dz2, uSSE = -0.25 * D23S * inS * inSEv * inSSE / dz2;
            const double uNNW = -0.25 * D23N * inN * inNNW * inNWv / dz2, uNNE = -0.25 * D23N * inN * inNEv * inNNE / dz2;
            const double pshear = (D34W * inW * wW - D34E * inE * wE) / dx;
            const double pS = -inS / dz + inS * (D24S / dz + pshear), pN = inN / dz + inN * (-D24N / dz + pshear);
            const double pSW = D34W * inSWc * inW * wW / dx, pSE = -D34E * inE * inSEc * wE / dx;
            const double pNW = D34W * inNWc * inW * wW / dx, pNE = -D34E * inE * inNEc * wE / dx;
            /* no-slip ghosts (:930-939) */
            if (Vt[iVzW] == 11) vC -= vW;
            if (Vt[iVzE] == 11) vC -= vE;
            if (Vt[iVzSW] == 11) vW -= vSW;
            if (Vt[iVzSE] == 11) vE -= vSE;
            if (Vt[iVzNW] == 11) vW -= vNW;
            if (Vt[iVzNE] == 11) vE -= vNE;
            if (Ut[iVxNNW] == 11) uNW -= uNNW;
            if (Ut[iVxNNE] == 11) uNE -= uNNE;
            if (Ut[iVxSSW] == 11) uSW -= uSSW;
            if (Ut[iVxSSE] == 11) uSE -= uSSE;
            /* entries in the reference's order (:961-1049) */
            if (Ut[iVxSSW] != 30) put(&A, Ut[iVxSSW], eqn_u[iVxSSW], uSSW * celvol);
            if (Ut[iVxSSE] != 30) put(&A, Ut[iVxSSE], eqn_u[iVxSSE], uSSE * celvol);
            if ((k > 1 || Ut[iVxSWW] == -1) && Ut[iVxSWW] != 30) put(&A, Ut[iVxSWW], eqn_u[iVxSWW], uSWW * celvol);
            if (Ut[iVxSW] != 30) put(&A, Ut[iVxSW], eqn_u[iVxSW], uSW * celvol);
            if (Ut[iVxSE] != 30) put(&A, Ut[iVxSE], eqn_u[iVxSE], uSE * celvol);
            if ((k < nx - 1 || Ut[iVxSEE] == -1) && Ut[c1 + 1] != 30) put(&A, Ut[iVxSEE], eqn_u[iVxSEE], uSEE * celvol);
            if ((k > 1 || Ut[iVxNWW] == -1) && Ut[c1 + nx - 2] != 30) put(&A, Ut[iVxNWW], eqn_u[iVxNWW], uNWW * celvol);
            if (Ut[iVxNW] != 30) put(&A, Ut[iVxNW], eqn_u[iVxNW], uNW * celvol);
            if (Ut[iVxNE] != 30) put(&A, Ut[iVxNE], eqn_u[iVxNE], uNE * celvol);
            if ((k < nx - 1 || Ut[iVxNEE] == -1) && Ut[iVxNEE] != 30) put(&A, Ut[iVxNEE], eqn_u[iVxNEE], uNEE * celvol);
            if (Ut[iVxNNW] != 30) put(&A, Ut[iVxNNW], eqn_u[iVxNNW], uNNW * celvol);
            if (Ut[iVxNNE] != 30) put(&A, Ut[iVxNNE], eqn_u[iVxNNE], uNNE * celvol);
            if (Vt[iVzSW] != 30 && l > 0) put(&A, Vt[iVzSW], eqn_v[iVzSW], vSW * celvol);
            if (Vt[iVzS] != 30 && l > 0) put(&A, Vt[iVzS], eqn_v[iVzS], vS * celvol);
            if (Vt[iVzSE] != 30 && l > 0) put(&A, Vt[iVzSE], eqn_v[iVzSE], vSE * celvol);
            if (Vt[iVzW] != 30) put(&A, Vt[iVzW], eqn_v[iVzW], vW * celvol);
            if (Vt[c3] != 30) put(&A, Vt[c3], eqn_v[c3], vC * celvol);
            if (Vt[iVzE] != 30) put(&A, Vt[iVzE], eqn_v[iVzE], vE * celvol);
            if (Vt[iVzNW] != 30 && l < nz - 1) put(&A, Vt[iVzNW], eqn_v[iVzNW], vNW * celvol);
            if (Vt[iVzN] != 30 && l < nz - 1) put(&A, Vt[iVzN], eqn_v[iVzN], vN * celvol);
            if (Vt[iVzNE] != 30 && l < nz - 1) put(&A, Vt[iVzNE], eqn_v[iVzNE], vNE * celvol);
            if (k > 1 && Pt[iPrSW] != 30) put(&B, Pt[iPrSW], eqn_p[iPrSW] - neq_mom, pSW * celvol);
            if (Pt[iPrS] != 30 && Pt[iPrN] != 30) put(&B, Pt[iPrS], eqn_p[iPrS] - neq_mom, pS * celvol);
            if (k < nx - 1 && Pt[iPrSE] != 30) put(&B, Pt[iPrSE], eqn_p[iPrSE] - neq_mom, pSE * celvol);
            if (k > 1 && Pt[iPrNW] != 30) put(&B, Pt[iPrNW], eqn_p[iPrNW] - neq_mom, pNW * celvol);
            if (Pt[iPrS] != 30 && Pt[iPrN] != 30) put(&B, Pt[iPrN], eqn_p[iPrN] - neq_mom, pN * celvol);
            if (k < nx - 1 && Pt[iPrNE] != 30) put(&B, Pt[iPrNE], eqn_p[iPrNE] - neq_mom, pNE * celvol);
        }
    Ap[row] = A.n;
    Bp[row] = B.n;
}
