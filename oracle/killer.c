/*
 * Oracle: KillerSolver (Solvers.c:1054-1454) and kspgcr (Solvers.c:124-230)
 * for the incompressible case.  TEST INFRASTRUCTURE ONLY.
 *
 * The reference preconditions GCR with a CHOLMOD (AMD-ordered supernodal)
 * factor of Jts = A - B*(penalty*I)*C.  SuiteSparse is not part of this
 * build; the factor here is an exact banded Cholesky factor of the same
 * matrix after a row-interleaved reordering (Vx row l, Vz row l, ...), which
 * is the same preconditioner up to factorisation round-off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "mdoracle.h"

typedef struct {
    int n, b;          /* size, half bandwidth                      */
    int *perm;         /* original index -> band index              */
    double *L;         /* row-major lower band, row i holds j in [i-b, i] */
} band_chol;

static void csr_mv(int n, const int *p, const int *j, const double *x, const double *v,
                   double *y, double alpha, double beta) {
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int q = p[i]; q < p[i + 1]; q++) s += x[q] * v[j[q]];
        y[i] = (beta == 0.0) ? alpha * s : beta * y[i] + alpha * s;   /* y may be uninitialised */
    }
}

/* Build the band factor of Jts = A - B*pen*C. */
static int band_factor(band_chol *bc, int nu, int np, const int *Ap, const int *Aj,
                       const double *Ax, const int *Bp, const int *Bj, const double *Bx,
                       const int *Cp, const int *Cj, const double *Cx, double pen,
                       const int *order) {
    bc->n = nu;
    bc->perm = (int *)malloc(sizeof(int) * nu);
    for (int i = 0; i < nu; i++) bc->perm[order[i]] = i;
    /* B transposed (columns of B as rows) */
    int *Tp = (int *)calloc(np + 1, sizeof(int));
    for (int q = 0; q < Bp[nu]; q++) Tp[Bj[q] + 1]++;
    for (int i = 0; i < np; i++) Tp[i + 1] += Tp[i];
    int *Tj = (int *)malloc(sizeof(int) * (Bp[nu] + 1));
    double *Tx = (double *)malloc(sizeof(double) * (Bp[nu] + 1));
    int *fill = (int *)malloc(sizeof(int) * (np + 1));
    memcpy(fill, Tp, sizeof(int) * (np + 1));
    for (int i = 0; i < nu; i++)
        for (int q = Bp[i]; q < Bp[i + 1]; q++) {
            Tj[fill[Bj[q]]] = i;
            Tx[fill[Bj[q]]++] = Bx[q];
        }
    free(fill);
    int b = 0;
    for (int i = 0; i < nu; i++)
        for (int q = Ap[i]; q < Ap[i + 1]; q++) {
            int d = abs(bc->perm[i] - bc->perm[Aj[q]]);
            if (d > b) b = d;
        }
    for (int p = 0; p < np; p++)
        for (int a = Tp[p]; a < Tp[p + 1]; a++)
            for (int c = Cp[p]; c < Cp[p + 1]; c++) {
                int d = abs(bc->perm[Tj[a]] - bc->perm[Cj[c]]);
                if (d > b) b = d;
            }
    bc->b = b;
    const int w = b + 1;
    double *L = (double *)calloc((size_t)nu * w, sizeof(double));
    bc->L = L;
#define BAND(i, j) L[(size_t)(i) * w + ((j) - (i) + b)]
    for (int i = 0; i < nu; i++)
        for (int q = Ap[i]; q < Ap[i + 1]; q++) {
            int r = bc->perm[i], c = bc->perm[Aj[q]];
            if (c <= r) BAND(r, c) += Ax[q];
        }
    for (int p = 0; p < np; p++)
        for (int a = Tp[p]; a < Tp[p + 1]; a++)
            for (int c = Cp[p]; c < Cp[p + 1]; c++) {
                int r = bc->perm[Tj[a]], cc = bc->perm[Cj[c]];
                if (cc <= r) BAND(r, cc) -= pen * Tx[a] * Cx[c];
            }
    free(Tp); free(Tj); free(Tx);
    /* in-place banded Cholesky, row oriented */
    for (int i = 0; i < nu; i++) {
        const int j0 = (i - b > 0) ? i - b : 0;
        for (int j = j0; j <= i; j++) {
            const int k0 = (j - b > j0) ? j - b : j0;
            double s = BAND(i, j);
            const double *li = &BAND(i, k0), *lj = &BAND(j, k0);
            for (int k = 0; k < j - k0; k++) s -= li[k] * lj[k];
            if (j == i) {
                if (!(s > 0.0)) return -1;
                BAND(i, i) = sqrt(s);
            } else {
                BAND(i, j) = s / BAND(j, j);
            }
        }
    }
#undef BAND
    return 0;
}

static void band_solve(const band_chol *bc, const double *rhs, double *x, double *tmp) {
    const int n = bc->n, b = bc->b, w = b + 1;
    const double *L = bc->L;
    for (int i = 0; i < n; i++) tmp[bc->perm[i]] = rhs[i];
    for (int i = 0; i < n; i++) {
        const int j0 = (i - b > 0) ? i - b : 0;
        double s = tmp[i];
        const double *li = L + (size_t)i * w + (j0 - i + b);
        for (int k = 0; k < i - j0; k++) s -= li[k] * tmp[j0 + k];
        tmp[i] = s / L[(size_t)i * w + b];
    }
    for (int i = n - 1; i >= 0; i--) {
        tmp[i] /= L[(size_t)i * w + b];
        const int j0 = (i - b > 0) ? i - b : 0;
        const double *li = L + (size_t)i * w + (j0 - i + b);
        for (int k = 0; k < i - j0; k++) tmp[j0 + k] -= li[k] * tmp[i];
    }
    for (int i = 0; i < n; i++) x[i] = tmp[bc->perm[i]];
}

typedef struct {
    int nu, np;
    const int *Ap, *Aj, *Bp, *Bj, *Cp, *Cj;
    const double *Ax, *Bx, *Cx;
    double pen;
    double *tp;   /* np scratch */
} js_op;

/* v = Js*s = (A - B*pen*C) s  (Solvers.c:1191, applied matrix-free here) */
static void js_apply(const js_op *o, const double *s, double *v) {
    csr_mv(o->nu, o->Ap, o->Aj, o->Ax, s, v, 1.0, 0.0);
    csr_mv(o->np, o->Cp, o->Cj, o->Cx, s, o->tp, 1.0, 0.0);
    for (int i = 0; i < o->np; i++) o->tp[i] *= o->pen;
    csr_mv(o->nu, o->Bp, o->Bj, o->Bx, o->tp, v, -1.0, 1.0);
}

static double dot(const double *a, const double *b, int n) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Solvers.c:124-230.  The orthogonalisation loop runs to i2 <= i1 with the
 * coefficient and stored pair of slot i1 left from the previous restart
 * cycle (zero in the first cycle); this is kept as in the reference. */
static int kspgcr(const js_op *M, const double *b, double *x, const band_chol *L, int N,
                  double eps, int max_it, double *work) {
    const int restart = 6;
    double *f = work, *v = f + N, *s = v + N, *tmp = s + N;
    double *VV = tmp + N, *SS = VV + (size_t)restart * N;
    double val[6] = {0, 0, 0, 0, 0, 0};
    memset(VV, 0, sizeof(double) * (size_t)restart * N * 2);
    js_apply(M, x, v);
    for (int i = 0; i < N; i++) f[i] = b[i] - v[i];
    double norm_r = sqrt(dot(f, f, N));
    const double rnorm0 = norm_r;
    int its = 0, success = 0;
    while (!success && its < max_it) {
        for (int i1 = 0; i1 < restart; i1++) {
            band_solve(L, f, s, tmp);
            js_apply(M, s, v);
            for (int i2 = 0; i2 < i1; i2++) val[i2] = dot(v, VV + (size_t)i2 * N, N);
            for (int i2 = 0; i2 < i1 + 1; i2++) {
                const double fact = -val[i2];
                const double *vv = VV + (size_t)i2 * N, *ss = SS + (size_t)i2 * N;
                for (int i = 0; i < N; i++) { v[i] += fact * vv[i]; s[i] += fact * ss[i]; }
            }
            double r_dot_v = dot(f, v, N);
            const double nrm = sqrt(dot(v, v, N));
            r_dot_v /= nrm;
            for (int i = 0; i < N; i++) { v[i] /= nrm; s[i] /= nrm; }
            for (int i = 0; i < N; i++) { x[i] += r_dot_v * s[i]; f[i] -= r_dot_v * v[i]; }
            norm_r = sqrt(dot(f, f, N));
            if (norm_r < eps * rnorm0) { success = 1; break; }
            memcpy(VV + (size_t)i1 * N, v, sizeof(double) * N);
            memcpy(SS + (size_t)i1 * N, s, sizeof(double) * N);
            its++;
        }
        its++;
    }
    return its;
}

/* MiscFunctions.c:342-375: signed maximum. */
static double signed_max(const double *a, int n) {
    double m = a[0];
    for (int i = 1; i < n; i++) if (a[i] > m) m = a[i];
    return m;
}

/* KillerSolver with the Picard blocks A, B (preconditioner Jts = A - B pen C,
 * Solvers.c:1150-1243 with preconditioner = 0) and the blocks JA, JB of the
 * operator GCR iterates on and the residuals are taken with (the Jacobian
 * blocks AcmJ, BcmJ, Solvers.c:1176-1193,1276-1314; A, B themselves for a
 * Picard solve). */
static int killer_core(int nu, int np, const int *Ap, const int *Aj, const double *Ax,
                       const int *Bp, const int *Bj, const double *Bx, const int *JAp, const int *JAj,
                       const double *JAx, const int *JBp, const int *JBj, const double *JBx,
                       const int *Cp, const int *Cj, const double *Cx, const double *bu, const double *bp,
                       double *du, double *dp, orc_killer_opts *o) {
    /* The band order interleaves rows: callers pass DOFs numbered Vx rows
     * then Vz rows; the row-interleaved order is supplied by
     * orc_band_order() through opts-independent numbering below. */
    extern const int *orc_band_order_current(void);
    const int *order = orc_band_order_current();
    int *ident = NULL;
    if (!order) {
        ident = (int *)malloc(sizeof(int) * nu);
        for (int i = 0; i < nu; i++) ident[i] = i;
        order = ident;
    }
    const double pen = o->penalty / o->celvol;   /* Solvers.c:1090 */
    band_chol L;
    o->status = 0;
    if (band_factor(&L, nu, np, Ap, Aj, Ax, Bp, Bj, Bx, Cp, Cj, Cx, pen, order) != 0) {
        o->status = -1;   /* Solvers.c:1246-1251 (not positive definite) */
        free(L.perm); free(L.L); free(ident);
        return -1;
    }
    double *work = (double *)calloc((size_t)nu * 16, sizeof(double));  /* kspgcr */
    double *tp = (double *)calloc(np, sizeof(double));
    double *ud = (double *)calloc(nu, sizeof(double)), *fu = (double *)calloc(nu, sizeof(double));
    double *pd = (double *)calloc(np, sizeof(double)), *fpv = (double *)calloc(np, sizeof(double));
    js_op M = {nu, np, JAp, JAj, JBp, JBj, Cp, Cj, JAx, JBx, Cx, pen, tp};
    memset(du, 0, sizeof(double) * nu);
    memset(dp, 0, sizeof(double) * np);
    double maxdiv = 0, maxdiv0 = 1, maxru = 0, maxru0 = 1;
    int its_tot = 0, k;
    for (k = 0; k < o->max_its_PH; k++) {
        /* udum = bu - B*(pen*bp) - B*dp  (Solvers.c:1290-1294) */
        for (int i = 0; i < np; i++) pd[i] = pen * bp[i] + dp[i];
        memcpy(ud, bu, sizeof(double) * nu);
        csr_mv(nu, JBp, JBj, JBx, pd, ud, -1.0, 1.0);
        its_tot += kspgcr(&M, ud, du, &L, nu, o->rel_tol_KSP, o->max_its_KSP, work);
        /* dp += pen*(bp - C*du)  (Solvers.c:1302-1305) */
        csr_mv(np, Cp, Cj, Cx, du, pd, 1.0, 0.0);
        for (int i = 0; i < np; i++) dp[i] += pen * (bp[i] - pd[i]);
        /* fu = bu - A du - B dp ; fp = bp - C du */
        memcpy(fu, bu, sizeof(double) * nu);
        csr_mv(nu, JAp, JAj, JAx, du, fu, -1.0, 1.0);
        csr_mv(nu, JBp, JBj, JBx, dp, fu, -1.0, 1.0);
        csr_mv(np, Cp, Cj, Cx, du, fpv, 1.0, 0.0);
        for (int i = 0; i < np; i++) fpv[i] = bp[i] - fpv[i];
        maxdiv = signed_max(fpv, np);
        maxru = signed_max(fu, nu);
        if (k == 0) { maxdiv0 = maxdiv; maxru0 = maxru; }
        if (k > 1 && (fabs(maxdiv) < o->lin_abs_div || maxdiv / maxdiv0 < o->lin_rel_div) &&
            (fabs(maxru) < o->lin_abs_mom || maxru / maxru0 < o->lin_rel_mom))
            break;
    }
    o->its_PH = (k < o->max_its_PH) ? k + 1 : k;
    o->its_KSP_tot = its_tot;
    o->maxdiv = maxdiv; o->maxdiv0 = maxdiv0; o->maxru = maxru; o->maxru0 = maxru0;
    /* Solvers.c:1333-1336 */
    if (fabs(maxdiv) > 10.0 * o->lin_abs_div && maxdiv / maxdiv0 > o->lin_rel_div) o->status = -2;
    for (int i = 0; i < nu; i++)
        if (isnan(du[i])) { o->status = -3; break; }
    free(work); free(tp); free(ud); free(fu); free(pd); free(fpv);
    free(L.perm); free(L.L); free(ident);
    return o->status;
}

int orc_killer_solve(int nu, int np, const int *Ap, const int *Aj, const double *Ax,
                     const int *Bp, const int *Bj, const double *Bx, const int *Cp,
                     const int *Cj, const double *Cx, const double *bu, const double *bp,
                     double *du, double *dp, orc_killer_opts *o) {
    return killer_core(nu, np, Ap, Aj, Ax, Bp, Bj, Bx, Ap, Aj, Ax, Bp, Bj, Bx, Cp, Cj, Cx, bu, bp, du, dp, o);
}

int orc_killer_solve_newton(int nu, int np, const int *Ap, const int *Aj, const double *Ax,
                            const int *Bp, const int *Bj, const double *Bx, const int *JAp,
                            const int *JAj, const double *JAx, const int *JBp, const int *JBj,
                            const double *JBx, const int *Cp, const int *Cj, const double *Cx,
                            const double *bu, const double *bp, double *du, double *dp,
                            orc_killer_opts *o) {
    return killer_core(nu, np, Ap, Aj, Ax, Bp, Bj, Bx, JAp, JAj, JAx, JBp, JBj, JBx, Cp, Cj, Cx, bu, bp, du, dp,
                       o);
}

/* Row-interleaved DOF order used for the band factor: for the momentum
 * numbering of StokesRoutines.c:1162-1220 (all Vx rows, then all Vz rows),
 * visit Vx row l, then Vz row l, for l = 0, 1, ... */
static int *g_order = NULL;
const int *orc_band_order_current(void) { return g_order; }

void orc_set_band_order(const orc_grid *g, const int *eqn_u, const int *eqn_v, int neq_mom) {
    free(g_order);
    g_order = (int *)malloc(sizeof(int) * neq_mom);
    const int nx = g->Nx, nz = g->Nz;
    int m = 0;
    for (int l = 0; l < nz + 1; l++) {
        for (int k = 0; k < nx; k++) {
            const int e = eqn_u[k + l * nx];
            if (e >= 0 && (g->BCu_t[k + l * nx] != -12)) g_order[m++] = e;
        }
        if (l < nz)
            for (int k = 0; k < nx + 1; k++) {
                const int e = eqn_v[k + l * (nx + 1)];
                if (e >= 0) g_order[m++] = e;
            }
    }
}
