// Newton linearisation: tangent coefficients and the matrix-free Jacobian.
//
// Tangent (RheologicalOperators with Jacobian = 1, StokesRoutines.c:250-375,
// isotropic and incompressible): the derivatives of the phase-averaged
// effective viscosity w.r.t. Exx, Ezz, gxz = 2 Exz and P are central finite
// differences of ViscosityConcise (DerivativesOnTheFly_n/_s and
// PhaseRheologyLoop_v1, FD_Jacobian.c:48-323), eight extra evaluations per
// active phase and node; then
//   D11 = 2 eta + 2 eta_xx Exx, D12 = 2 eta_zz Exx, D13 = 2 eta_g Exx, D14 = 2 eta_p Exx
//   D21 = 2 eta_xx Ezz, D22 = 2 eta + 2 eta_zz Ezz, D23 = 2 eta_g Ezz, D24 = 2 eta_p Ezz
//   D31 = 2 eta_xx Exz, D32 = 2 eta_zz Exz, D33 = eta + 2 eta_g Exz, D34 = 2 eta_p Exz.
//
// Jacobian-vector product: J v is the linearised residual
// (StokesAssemblyDecoupled.c:219-1085 assembles the same operator): the
// strain rates of the correction (ghost rows and columns from the
// homogeneous boundary relations), centre and vertex averages as in the
// residual path (exz to the centres, exx/ezz/p to the vertices), tangent
// stresses
//   dsxx = D11 exx + D12 ezz + D13 gxz_c + D14 p  (centres), dszz likewise,
//   dsxz = D31 exx_v + D32 ezz_v + D33 gxz + D34 p_v  (vertices),
// and the momentum/continuity rows of the residual on those stresses.
#include "mdg_internal.h"
#include "rheo_device.h"
#include "stokes_internal.h"

using rheo::rheo_ctl;
using rheo::viscosity_concise;

namespace {

__global__ void __launch_bounds__(256)
k_tangent_grid(rheo_ctl md, const mdg_phase_prop *__restrict__ phases, int n,
               const signed char *__restrict__ tag, int centroid, mdg_rheo_nodes f,
               double *__restrict__ D, int *__restrict__ fail) {
    __shared__ mdg_phase_prop sph[MDG_MAX_PHASES];
    for (int p = threadIdx.x; p < md.nph; p += blockDim.x) sph[p] = phases[p];
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int nrow = centroid ? 8 : 4;
    const signed char t = tag[i];
    if (t == 30 || (centroid && t == 31)) {
        for (int r = 0; r < nrow; r++) D[(size_t)r * n + i] = 0.0;
        return;
    }
    const double el = md.elastic ? 1.0 : 0.0;
    const double G = f.G[i];
    const double eta_e = md.elastic ? md.dt * G : 1.0;
    // EffectiveStrainRate (RheologyDensity.c:116-146), isotropic
    const double Exx = f.exx[i] + (f.sxx0 ? el * f.sxx0[i] / eta_e / 2.0 : 0.0);
    const double Ezz = f.ezz[i] + (f.szz0 ? el * f.szz0[i] / eta_e / 2.0 : 0.0);
    const double Exz = f.exz[i] + (f.sxz0 ? el * f.sxz0[i] / eta_e / 2.0 : 0.0);
    const double T = f.T[i], P = f.P[i];
    const double pref = f.pref ? f.pref[i] : 1.0;
    const double tol = centroid ? 1e-5 : 1e-7;
    const double px = tol * fabs(Exx) + tol / 1e10, pz = tol * fabs(Ezz) + tol / 1e10;
    const double pg = tol * fabs(Exz) + tol / 1e10, pp = tol * fabs(P) + tol / 1e10;
    bool ok = true;
    double e = 0.0;
    for (int p = 0; p < md.nph; p++) {
        const double w = f.perc[(size_t)p * n + i];
        if (!(fabs(w) > 1e-13)) continue;
        const mdg_phase_prop &m = sph[p];
        double etaVE, Pc, Epl;
        viscosity_concise(m, md, G, T, P, Exx, Ezz, Exz, f.fric ? f.fric[i] : m.fric,
                          f.C ? f.C[i] : m.C, f.dil ? f.dil[i] : m.dil, pref, etaVE, Pc, Epl, ok);
        if (md.average == 0) e += w * etaVE;
        else if (md.average == 1) e += w / etaVE;
        else e += w * log(etaVE);
    }
    if (md.average == 1) e = 1.0 / e;
    else if (md.average == 2) e = exp(e);
    // Central differences per phase: each phase's +h and -h values are
    // differenced before the 1/(2h) scaling (the reference sums the scaled
    // +h and -h terms of all phases, FD_Jacobian.c:66-78, leaving round-off
    // of size ulp(eta)/h in two-phase cells: O(1) at the h = 1e-15 floor of
    // a zero strain-rate component).  Same central difference, no cancellation.
    double der[4] = {0.0, 0.0, 0.0, 0.0};
    for (int q = 0; q < 4; q++) {
        const double h = (q == 0) ? px : (q == 1) ? pz : (q == 2) ? pg : pp;
        const double den = (q == 2) ? 4.0 * h : 2.0 * h;
        for (int p = 0; p < md.nph; p++) {
            const double w = f.perc[(size_t)p * n + i];
            if (!(fabs(w) > 1e-13)) continue;
            const mdg_phase_prop &m = sph[p];
            const double fr = f.fric ? f.fric[i] : m.fric, co = f.C ? f.C[i] : m.C, di = f.dil ? f.dil[i] : m.dil;
            double eb = 0.0, Pc, Epl;
            if (md.average != 0)   // the phase's own value at the base state
                viscosity_concise(m, md, G, T, P, Exx, Ezz, Exz, fr, co, di, pref, eb, Pc, Epl, ok);
            double epm[2];
            for (int sgn = 0; sgn < 2; sgn++) {
                const double d = sgn ? -h : h;
                viscosity_concise(m, md, G, T, P + (q == 3 ? d : 0.0), Exx + (q == 0 ? d : 0.0),
                                  Ezz + (q == 1 ? d : 0.0), Exz + (q == 2 ? d : 0.0), fr, co, di, pref, epm[sgn],
                                  Pc, Epl, ok);
            }
            double v = w * (epm[0] - epm[1]) / den;
            if (md.average == 1) v /= eb * eb;
            else if (md.average == 2) v /= eb;
            der[q] += v;
        }
    }
    for (int q = 0; q < 4; q++) {
        if (md.average == 1) der[q] *= e * e;
        else if (md.average == 2) der[q] *= e;
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
        D[0 * (size_t)n + i] = 2.0 * der[0] * Exz;
        D[1 * (size_t)n + i] = 2.0 * der[1] * Exz;
        D[2 * (size_t)n + i] = e + 2.0 * der[2] * Exz;
        D[3 * (size_t)n + i] = 2.0 * der[3] * Exz;
    }
    if (!ok) atomicOr(fail, 1);
}

// Correction velocities with ghost rows/columns from the homogeneous
// boundary relations: free slip copies the interior value, no slip negates it;
// Dirichlet walls stay zero.  x is the grid-layout vector [u | v | p].
__global__ void k_hom_ghosts(sgrid g, const double *__restrict__ x, double *__restrict__ u,
                             double *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Nz = g.Nz, nu = nvx(g), nv = nvz(g);
    if (i < nu) {
        const int k = i % Nx, l = i / Nx;
        double val = vx_free(g, k, l) ? x[i] : 0.0;
        if ((l == 0 || l == Nz) && k > 0 && k < Nx - 1) {
            const int ln = (l == 0) ? 1 : Nz - 1;
            const double in = x[k + ln * Nx];
            const int kind = (l == 0) ? g.ghostS : g.ghostN;
            val = (kind == 2) ? -in : in;
        }
        u[i] = val;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % (Nx + 1), l = j / (Nx + 1);
        double val = vz_free(g, k, l) ? x[nu + j] : 0.0;
        if ((k == 0 || k == Nx) && l > 0 && l < Nz - 1) {
            const int kn = (k == 0) ? 1 : Nx - 1;
            const double in = x[nu + kn + l * (Nx + 1)];
            const int kind = (k == 0) ? g.ghostW : g.ghostE;
            val = (kind == 2) ? -in : in;
        }
        v[j] = val;
    }
}

// Strain rates of the correction: div, exx on centres, gxz = 2 exz on vertices.
__global__ void k_jv_strain(sgrid g, const double *__restrict__ u, const double *__restrict__ v,
                            double *__restrict__ exx, double *__restrict__ ezz, double *__restrict__ gxz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Ncx = Nx - 1, nc = ncell(g), nvt = g.Nx * g.Nz;
    if (i < nc) {
        const int k = i % Ncx, l = i / Ncx;
        const int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
        exx[i] = (u[c1 + 1 + Nx] - u[c1 + Nx]) / g.dx;
        ezz[i] = (v[c2 + Nx + 1 + 1] - v[c2 + 1]) / g.dz;
        return;
    }
    i -= nc;
    if (i < nvt) {
        const int k = i % Nx, l = i / Nx;
        const int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
        gxz[i] = 2.0 * (0.5 * ((u[c1 + Nx] - u[c1]) / g.dz + (v[c2 + 1] - v[c2]) / g.dx));
    }
}

// Tangent stresses: centres from the centre rates and the 4-vertex mean of
// gxz; vertices from the mean of the surrounding centres (edge copy as in
// InterpCentroidsToVerticesDouble for a box without air) and gxz.
__global__ void k_jv_stress(sgrid g, const double *__restrict__ exx, const double *__restrict__ ezz,
                            const double *__restrict__ gxz, const double *__restrict__ p,
                            const double *__restrict__ Dn, const double *__restrict__ Ds,
                            double *__restrict__ sxx, double *__restrict__ szz, double *__restrict__ sxz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Nz = g.Nz, Ncx = Nx - 1, Ncz = Nz - 1, nc = ncell(g), nvt = Nx * Nz;
    if (i < nc) {
        const int k = i % Ncx, l = i / Ncx, v0 = k + l * Nx;
        const double gc = 0.25 * (gxz[v0] + gxz[v0 + 1] + gxz[v0 + Nx] + gxz[v0 + Nx + 1]);
        const double a = exx[i], b = ezz[i], pp = p[i];
        sxx[i] = Dn[0 * (size_t)nc + i] * a + Dn[1 * (size_t)nc + i] * b + Dn[2 * (size_t)nc + i] * gc + Dn[3 * (size_t)nc + i] * pp;
        szz[i] = Dn[4 * (size_t)nc + i] * a + Dn[5 * (size_t)nc + i] * b + Dn[6 * (size_t)nc + i] * gc + Dn[7 * (size_t)nc + i] * pp;
        return;
    }
    i -= nc;
    if (i < nvt) {
        const int k = i % Nx, l = i / Nx;
        double ax = 0.0, az = 0.0, ap = 0.0;
        for (int dl = -1; dl <= 0; dl++) {
            const int lc = min(max(l + dl, 0), Ncz - 1);
            for (int dk = -1; dk <= 0; dk++) {
                const int kc = min(max(k + dk, 0), Ncx - 1), c = kc + lc * Ncx;
                ax += exx[c]; az += ezz[c]; ap += p[c];
            }
        }
        ax *= 0.25; az *= 0.25; ap *= 0.25;
        sxz[i] = Ds[0 * (size_t)nvt + i] * ax + Ds[1 * (size_t)nvt + i] * az + Ds[2 * (size_t)nvt + i] * gxz[i] +
                 Ds[3 * (size_t)nvt + i] * ap;
    }
}

// y = J_gamma x (grid layout): momentum rows -dV (d sxx/dx - dp/dx + d sxz/dz)
// plus the viscosity-weighted penalty term, continuity rows dV div.
__global__ void k_jv_rows(sgrid g, double gamma, const double *__restrict__ u, const double *__restrict__ v,
                          const double *__restrict__ p, const double *__restrict__ sxx,
                          const double *__restrict__ szz, const double *__restrict__ sxz,
                          const double *__restrict__ en, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Ncx = Nx - 1, nu = nvx(g), nv = nvz(g), nc = ncell(g), NxV = Nx + 1;
    const double dV = g.dx * g.dz;
    if (i < nu) {
        const int k = i % Nx, l = i / Nx;
        double f = 0.0;
        if (vx_free(g, k, l)) {
            const int cW = k - 1 + (l - 1) * Ncx, cE = cW + 1;
            f = (sxx[cE] - sxx[cW]) / g.dx - (p[cE] - p[cW]) / g.dx + (sxz[i] - sxz[i - Nx]) / g.dz;
            f *= -dV;
            if (gamma != 0.0) {
                const double divW = (u[i] - u[i - 1]) / g.dx + (v[k + l * NxV] - v[k + (l - 1) * NxV]) / g.dz;
                const double divE = (u[i + 1] - u[i]) / g.dx + (v[k + 1 + l * NxV] - v[k + 1 + (l - 1) * NxV]) / g.dz;
                f += gamma * (en[cW] * divW - en[cE] * divE) / g.dx * dV;
            }
        }
        y[i] = f;
    } else if (i < nu + nv) {
        const int j = i - nu, k = j % NxV, l = j / NxV;
        double f = 0.0;
        if (vz_free(g, k, l)) {
            const int cS = k - 1 + (l - 1) * Ncx, cN = cS + Ncx, c1 = k + l * Nx;
            f = (szz[cN] - szz[cS]) / g.dz - (p[cN] - p[cS]) / g.dz + (sxz[c1] - sxz[c1 - 1]) / g.dx;
            f *= -dV;
            if (gamma != 0.0) {
                const double divS = (u[k + l * Nx] - u[k - 1 + l * Nx]) / g.dx + (v[j] - v[j - NxV]) / g.dz;
                const double divN = (u[k + (l + 1) * Nx] - u[k - 1 + (l + 1) * Nx]) / g.dx + (v[j + NxV] - v[j]) / g.dz;
                f += gamma * (en[cS] * divS - en[cN] * divN) / g.dz * dV;
            }
        }
        y[i] = f;
    } else if (i < nu + nv + nc) {
        const int c = i - nu - nv, k = c % Ncx, l = c / Ncx;
        y[i] = dV * ((u[k + 1 + (l + 1) * Nx] - u[k + (l + 1) * Nx]) / g.dx +
                     (v[k + 1 + (l + 1) * NxV] - v[k + 1 + l * NxV]) / g.dz);
    }
}

}  // namespace

// Device state of the Newton operator (coefficients and work arrays).
struct newton_state {
    dbuf<double> Dn, Ds;                 // 8 x cells, 4 x vertices
    dbuf<double> u, v, exx, ezz, gxz, sxx, szz, sxz;
    dbuf<mdg_phase_prop> ph;
    bool active = false;
};

newton_state *newton_of(mdg_ctx *c) {
    if (!c->newton) c->newton = new newton_state;
    return c->newton;
}

void mdg_newton_release(mdg_ctx *c) {
    newton_state *s = c->newton;
    if (!s) return;
    s->Dn.release(); s->Ds.release(); s->u.release(); s->v.release(); s->exx.release();
    s->ezz.release(); s->gxz.release(); s->sxx.release(); s->szz.release(); s->sxz.release();
    s->ph.release();
    delete s;
    c->newton = nullptr;
}

// y = J_gamma x on grid-layout device vectors, with the coefficients staged by
// mdg_newton_set; gamma = 0 gives J itself.
int mdg_newton_apply_grid(mdg_ctx *c, double gamma, const double *x, double *y) {
    newton_state *s = c->newton;
    const sgrid g = c->sg;
    const int nu = nvx(g), nv = nvz(g), nc = ncell(g), nvt = g.Nx * g.Nz;
    hipLaunchKernelGGL(k_hom_ghosts, dim3(mdg_grid1d(nu + nv, 256)), dim3(256), 0, c->stream, g, x, s->u.p, s->v.p);
    hipLaunchKernelGGL(k_jv_strain, dim3(mdg_grid1d(nc + nvt, 256)), dim3(256), 0, c->stream, g, (const double *)s->u.p,
                       (const double *)s->v.p, s->exx.p, s->ezz.p, s->gxz.p);
    hipLaunchKernelGGL(k_jv_stress, dim3(mdg_grid1d(nc + nvt, 256)), dim3(256), 0, c->stream, g, (const double *)s->exx.p,
                       (const double *)s->ezz.p, (const double *)s->gxz.p, x + nu + nv, (const double *)s->Dn.p,
                       (const double *)s->Ds.p, s->sxx.p, s->szz.p, s->sxz.p);
    hipLaunchKernelGGL(k_jv_rows, dim3(mdg_grid1d(nu + nv + nc, 256)), dim3(256), 0, c->stream, g, gamma,
                       (const double *)s->u.p, (const double *)s->v.p, x + nu + nv, (const double *)s->sxx.p,
                       (const double *)s->szz.p, (const double *)s->sxz.p, (const double *)c->eta_n.p, y);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

bool mdg_newton_active(const mdg_ctx *c) { return c->newton && c->newton->active; }

extern "C" int mdg_tangent_grid(mdg_ctx *c, const mdg_phase_prop *phases, int nph,
                                const mdg_rheo_params *prm, int centroid, const mdg_rheo_nodes *f,
                                double *D) {
    if (!c || !phases || !prm || !f || !D) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (nph < 1 || nph > MDG_MAX_PHASES) return mdg_fail(c, MDG_ERR_ARG, "nph = %d outside [1, %d]", nph, MDG_MAX_PHASES);
    if (!f->perc || !f->G || !f->T || !f->P || !f->exx || !f->ezz || !f->exz)
        return mdg_fail(c, MDG_ERR_ARG, "perc, G, T, P, exx, ezz and exz are required");
    for (int p = 0; p < nph; p++)
        if (phases[p].plast != 0 && phases[p].plast != 1)
            return mdg_fail(c, MDG_ERR_UNSUPPORTED, "phase %d: plast = %d (only 0 and 1 are implemented)", p, phases[p].plast);
    const int n = centroid ? (c->Nx - 1) * (c->Nz - 1) : c->Nx * c->Nz;
    const int nrow = centroid ? 8 : 4;
    newton_state *s = newton_of(c);
    // stage inputs through the rheology scratch of the grid update
    std::vector<dbuf<double>> tmp(14);
    mdg_rheo_nodes d = {};
    const double *const *ins[14] = {&f->perc, &f->G, &f->T, &f->P, &f->exx, &f->ezz, &f->exz,
                                    &f->sxx0, &f->szz0, &f->sxz0, &f->fric, &f->C, &f->dil, &f->pref};
    const double **din[14] = {&d.perc, &d.G, &d.T, &d.P, &d.exx, &d.ezz, &d.exz,
                              &d.sxx0, &d.szz0, &d.sxz0, &d.fric, &d.C, &d.dil, &d.pref};
    int rc = MDG_OK;
    for (int q = 0; q < 14 && !rc; q++)
        rc = mdg_stage_in(c, *ins[q], (size_t)n * (q == 0 ? nph : 1), tmp[q], din[q]);
    dbuf<double> dout;
    double *Dd = D;
    if (!rc && !mdg_is_device_ptr(D)) {
        if (dout.reserve((size_t)nrow * n) != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "allocation failed");
        Dd = dout.p;
    }
    if (!rc && s->ph.reserve(nph) != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "allocation failed");
    int fail = 0;
    if (!rc) {
        rheo_ctl md = {prm->step, prm->elastic, prm->average, nph, prm->dt, prm->R, prm->zeroC,
                       prm->min_eta, prm->max_eta, prm->eta_tol, prm->Tiimin, prm->scale_E, prm->scale_eta};
        if (hipMemcpyAsync(s->ph.p, phases, nph * sizeof(mdg_phase_prop), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            c->flag_dev.reserve(4) != hipSuccess ||
            hipMemsetAsync(c->flag_dev.p, 0, sizeof(int), c->stream) != hipSuccess)
            rc = mdg_fail(c, MDG_ERR_HIP, "staging failed");
        else {
            hipLaunchKernelGGL(k_tangent_grid, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, md,
                               (const mdg_phase_prop *)s->ph.p, n, centroid ? c->BCp_t.p : c->BCg_t.p,
                               centroid ? 1 : 0, d, Dd, c->flag_dev.p);
            if (hipGetLastError() != hipSuccess) rc = mdg_fail(c, MDG_ERR_HIP, "tangent kernel launch failed");
        }
        if (!rc && Dd != D &&
            hipMemcpyAsync(D, Dd, (size_t)nrow * n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            rc = mdg_fail(c, MDG_ERR_HIP, "copy back failed");
        if (!rc && (hipMemcpyAsync(&fail, c->flag_dev.p, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                    hipStreamSynchronize(c->stream) != hipSuccess))
            rc = mdg_fail(c, MDG_ERR_HIP, "synchronisation failed");
    }
    (void)hipStreamSynchronize(c->stream);
    for (auto &b : tmp) b.release();
    dout.release();
    if (rc) return rc;
    if (fail) return mdg_fail(c, MDG_ERR_DIVERGED, "visco-elastic local iterations failed to converge");
    return MDG_OK;
}

// Stage the Newton coefficients for mdg_jacobian_apply / the Newton solve.
static int newton_set(mdg_ctx *c, const double *Dn, const double *Ds) {
    newton_state *s = newton_of(c);
    const sgrid g = c->sg;
    const size_t nc = ncell(g), nvt = (size_t)g.Nx * g.Nz, nu = nvx(g), nv = nvz(g);
    HIPCHK(c, s->Dn.reserve(8 * nc));
    HIPCHK(c, s->Ds.reserve(4 * nvt));
    HIPCHK(c, s->u.reserve(nu)); HIPCHK(c, s->v.reserve(nv));
    HIPCHK(c, s->exx.reserve(nc)); HIPCHK(c, s->ezz.reserve(nc)); HIPCHK(c, s->gxz.reserve(nvt));
    HIPCHK(c, s->sxx.reserve(nc)); HIPCHK(c, s->szz.reserve(nc)); HIPCHK(c, s->sxz.reserve(nvt));
    HIPCHK(c, hipMemcpyAsync(s->Dn.p, Dn, 8 * nc * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipMemcpyAsync(s->Ds.p, Ds, 4 * nvt * 8, hipMemcpyDefault, c->stream));
    return MDG_OK;
}

extern "C" int mdg_jacobian_apply(mdg_ctx *c, const double *Dn, const double *Ds, const double *x, double *y) {
    if (!c || !Dn || !Ds || !x || !y) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc;
    if ((rc = newton_set(c, Dn, Ds)) || (rc = mdg_reserve_grid_vectors(c))) return rc;
    HIPCHK(c, c->eta_n.reserve(ncell(c->sg)));
    const sgrid g = c->sg;
    const double *xd;
    if ((rc = mdg_stage_in(c, x, (size_t)c->neq, c->tmp_c, &xd))) return rc;
    double *gxv = c->gx.p, *gyv = c->gy.p;
    if ((rc = mdg_eqn_to_grid(c, xd, gxv, gxv + nvx(g), gxv + nvx(g) + nvz(g), 1.0))) return rc;
    if ((rc = mdg_newton_apply_grid(c, 0.0, gxv, gyv))) return rc;
    double *yd = mdg_is_device_ptr(y) ? y : c->eqbuf.p;
    if ((rc = mdg_grid_to_eqn(c, gyv, gyv + nvx(g), gyv + nvx(g) + nvz(g), yd))) return rc;
    if (yd != y) HIPCHK(c, hipMemcpyAsync(y, yd, (size_t)c->neq * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_newton_solver(mdg_ctx *c, mdg_solver_opts *o, const double *Dn, const double *Ds,
                                 const double *D11_n, const double *D33_s, const double *rhs_mom,
                                 const double *rhs_cont, double *sol) {
    if (!c || !Dn || !Ds) return mdg_fail(c, MDG_ERR_ARG, "null argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc = newton_set(c, Dn, Ds);
    if (rc) return rc;
    c->newton->active = true;
    rc = mdg_killer_solver(c, o, D11_n, D11_n, D33_s, rhs_mom, rhs_cont, sol);
    c->newton->active = false;
    return rc;
}
