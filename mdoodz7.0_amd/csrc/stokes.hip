// Stokes operator, residual and strain-rate kernels (FP64) and the DOF
// numbering of the reference (StokesRoutines.c:1162-1220).
#include <mutex>

#include "stokes_internal.h"

// ------------------------------------------------------------------ numbering
// Vx rows, then Vz rows, then pressure, row-major with k fastest.
int mdg_build_numbering(mdg_ctx *c, const std::vector<signed char> &tu,
                        const std::vector<signed char> &tv, const std::vector<signed char> &tp) {
    const int Nx = c->Nx, Nz = c->Nz;
    std::vector<int> eu(tu.size(), -1), ev(tv.size(), -1), ep(tp.size(), -1);
    int inc = 0;
    for (int l = 0; l < Nz + 1; l++)
        for (int k = 0; k < Nx; k++) {
            const int i = k + l * Nx;
            if (tu[i] == -1 || tu[i] == 2 || tu[i] == -2) eu[i] = inc++;
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx + 1; k++) {
            const int i = k + l * (Nx + 1);
            if (tv[i] == -1 || tv[i] == 2) ev[i] = inc++;
        }
    c->neq_mom = inc;
    // rows entering the residual norms (StokesRoutines.c:890-916)
    c->ndof_norm[0] = c->ndof_norm[1] = 0;
    for (int l = 0; l < Nz + 1; l++)
        for (int k = 0; k < Nx; k++)
            if (eu[k + l * Nx] >= 0 && l > 2 && l < Nz + 1 - 2) c->ndof_norm[0]++;
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx + 1; k++)
            if (ev[k + l * (Nx + 1)] >= 0 && k > 2 && k < Nx + 1 - 2) c->ndof_norm[1]++;
    for (size_t i = 0; i < tp.size(); i++)
        if (tp[i] == -1) ep[i] = inc++;
    c->neq = inc;
    c->neq_cont = inc - c->neq_mom;
    HIPCHK(c, c->eqn_u.reserve(eu.size()));
    HIPCHK(c, c->eqn_v.reserve(ev.size()));
    HIPCHK(c, c->eqn_p.reserve(ep.size()));
    HIPCHK(c, hipMemcpy(c->eqn_u.p, eu.data(), eu.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->eqn_v.p, ev.data(), ev.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->eqn_p.p, ep.data(), ep.size() * 4, hipMemcpyHostToDevice));
    return MDG_OK;
}

// equation vector <-> grid arrays (zeros where there is no unknown)
__global__ void k_eqn_to_grid(int n, const int *__restrict__ eqn, const double *__restrict__ x,
                              double *__restrict__ g, double scale) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = eqn[i];
    g[i] = (e >= 0) ? scale * x[e] : 0.0;
}
__global__ void k_grid_to_eqn(int n, const int *__restrict__ eqn, const double *__restrict__ g,
                              double *__restrict__ x) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int e = eqn[i];
    if (e >= 0) x[e] = g[i];
}

int mdg_eqn_to_grid(mdg_ctx *c, const double *x, double *u, double *v, double *p, double scale) {
    const sgrid g = c->sg;
    hipLaunchKernelGGL(k_eqn_to_grid, dim3(mdg_grid1d(nvx(g), 256)), dim3(256), 0, c->stream, nvx(g), c->eqn_u.p, x, u, scale);
    hipLaunchKernelGGL(k_eqn_to_grid, dim3(mdg_grid1d(nvz(g), 256)), dim3(256), 0, c->stream, nvz(g), c->eqn_v.p, x, v, scale);
    hipLaunchKernelGGL(k_eqn_to_grid, dim3(mdg_grid1d(ncell(g), 256)), dim3(256), 0, c->stream, ncell(g), c->eqn_p.p, x, p, scale);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}
int mdg_grid_to_eqn(mdg_ctx *c, const double *u, const double *v, const double *p, double *x) {
    const sgrid g = c->sg;
    hipLaunchKernelGGL(k_grid_to_eqn, dim3(mdg_grid1d(nvx(g), 256)), dim3(256), 0, c->stream, nvx(g), c->eqn_u.p, u, x);
    hipLaunchKernelGGL(k_grid_to_eqn, dim3(mdg_grid1d(nvz(g), 256)), dim3(256), 0, c->stream, nvz(g), c->eqn_v.p, v, x);
    hipLaunchKernelGGL(k_grid_to_eqn, dim3(mdg_grid1d(ncell(g), 256)), dim3(256), 0, c->stream, ncell(g), c->eqn_p.p, p, x);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

// ------------------------------------------------------------- full operator
// y = K x for the Picard Stokes operator K = [A B; C 0] (grid layout, FP64).
// One launch covers the Vx rows, the Vz rows and the cells.
__global__ void __launch_bounds__(256) k_apply_K(sgrid g, const double *__restrict__ u,
                                                 const double *__restrict__ v,
                                                 const double *__restrict__ p,
                                                 const double *__restrict__ en,
                                                 const double *__restrict__ es,
                                                 double *__restrict__ yu, double *__restrict__ yv,
                                                 double *__restrict__ yp) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nv = nvz(g), np = ncell(g);
    if (i < nu) {
        const int k = i % g.Nx, l = i / g.Nx;
        double y = 0.0;
        if (vx_free(g, k, l))
            y = apply_vx_row<double>(g, k, l, u, v, en, es, 0.0, nullptr) + grad_vx_row<double>(g, k, l, p);
        yu[i] = y;
        return;
    }
    i -= nu;
    if (i < nv) {
        const int k = i % (g.Nx + 1), l = i / (g.Nx + 1);
        double y = 0.0;
        if (vz_free(g, k, l))
            y = apply_vz_row<double>(g, k, l, u, v, en, es, 0.0, nullptr) + grad_vz_row<double>(g, k, l, p);
        yv[i] = y;
        return;
    }
    i -= nv;
    if (i < np) {
        const int k = i % (g.Nx - 1), l = i / (g.Nx - 1);
        yp[i] = div_cell<double>(g, k, l, u, v);
    }
}

int mdg_apply_K(mdg_ctx *c, const double *u, const double *v, const double *p, const double *en,
                const double *es, double *yu, double *yv, double *yp) {
    const sgrid g = c->sg;
    const int n = nvx(g) + nvz(g) + ncell(g);
    hipLaunchKernelGGL(k_apply_K, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, u, v, p, en,
                       es, yu, yv, yp);
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

// eta_n = D11/2 (isotropic Picard coefficients, StokesRoutines.c:113-122,195-198)
__global__ void k_half(int n, const double *__restrict__ a, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = 0.5 * a[i];
}

// Stage the Picard coefficients D11_n, D22_n, D33_s as eta_n, eta_s on the
// device.  D22 must equal D11 (isotropic rheology).
int mdg_stage_coeffs(mdg_ctx *c, const double *D11, const double *D22, const double *D33) {
    const sgrid g = c->sg;
    const int nn = ncell(g), ns = g.Nx * g.Nz;
    if (!D11 || !D33) return mdg_fail(c, MDG_ERR_ARG, "null rheological coefficients");
    if (D22 && D22 != D11) {
        std::vector<double> a(nn), b(nn);
        HIPCHK(c, hipMemcpy(a.data(), D11, nn * 8, hipMemcpyDefault));
        HIPCHK(c, hipMemcpy(b.data(), D22, nn * 8, hipMemcpyDefault));
        for (int i = 0; i < nn; i++)
            if (a[i] != b[i]) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "D22_n != D11_n (anisotropic) is not implemented");
    }
    const double *d11, *d33;
    int rc;
    if ((rc = mdg_stage_in(c, D11, (size_t)nn, c->tmp_a, &d11)) || (rc = mdg_stage_in(c, D33, (size_t)ns, c->tmp_b, &d33)))
        return rc;
    HIPCHK(c, c->eta_n.reserve(nn));
    HIPCHK(c, c->eta_s.reserve(ns));
    hipLaunchKernelGGL(k_half, dim3(mdg_grid1d(nn, 256)), dim3(256), 0, c->stream, nn, d11, c->eta_n.p);
    HIPCHK(c, hipMemcpyAsync(c->eta_s.p, d33, ns * 8, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipGetLastError());
    return MDG_OK;
}

int mdg_reserve_grid_vectors(mdg_ctx *c) {
    const sgrid g = c->sg;
    const size_t n = (size_t)nvx(g) + nvz(g) + ncell(g);
    HIPCHK(c, c->gx.reserve(n));
    HIPCHK(c, c->gy.reserve(n));
    HIPCHK(c, c->eqbuf.reserve((size_t)c->neq + 1));
    return MDG_OK;
}

extern "C" int mdg_stokes_apply(mdg_ctx *c, const double *D11_n, const double *D22_n,
                                const double *D33_s, const double *x, double *y) {
    if (!c || !x || !y) return mdg_fail(c, MDG_ERR_ARG, "null vector");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc;
    if ((rc = mdg_stage_coeffs(c, D11_n, D22_n, D33_s)) || (rc = mdg_reserve_grid_vectors(c))) return rc;
    const sgrid g = c->sg;
    double *u = c->gx.p, *v = u + nvx(g), *p = v + nvz(g);
    double *yu = c->gy.p, *yv = yu + nvx(g), *yp = yv + nvz(g);
    const double *xd;
    if ((rc = mdg_stage_in(c, x, (size_t)c->neq, c->tmp_c, &xd))) return rc;
    if ((rc = mdg_eqn_to_grid(c, xd, u, v, p, 1.0)) ||
        (rc = mdg_apply_K(c, u, v, p, c->eta_n.p, c->eta_s.p, yu, yv, yp)))
        return rc;
    double *yd = mdg_is_device_ptr(y) ? y : c->eqbuf.p;
    if ((rc = mdg_grid_to_eqn(c, yu, yv, yp, yd))) return rc;
    if (yd != y) HIPCHK(c, hipMemcpyAsync(y, yd, (size_t)c->neq * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_eqn_count(const mdg_ctx *c, int *neq, int *neq_mom, int *neq_cont) {
    if (!c || !c->have_bc) return MDG_ERR_STATE;
    if (neq) *neq = c->neq;
    if (neq_mom) *neq_mom = c->neq_mom;
    if (neq_cont) *neq_cont = c->neq_cont;
    return MDG_OK;
}

// ------------------------------------------------------- residual and norms
// StrainRateComponents (RheologyDensity.c:2001-2097, incompressible, no air):
// div and deviatoric normal rates on cells, exz on vertices from the full
// (boundary-filled) velocity arrays.
__global__ void k_strain_rates(sgrid g, const double *__restrict__ u, const double *__restrict__ v,
                               double *__restrict__ div, double *__restrict__ exx,
                               double *__restrict__ ezz, double *__restrict__ exz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Ncx = Nx - 1, nc = ncell(g), nvtx = g.Nx * g.Nz;
    if (i < nc) {
        const int k = i % Ncx, l = i / Ncx;
        const int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
        const double dvxdx = (u[c1 + 1 + Nx] - u[c1 + Nx]) / g.dx;
        const double dvzdz = (v[c2 + Nx + 1 + 1] - v[c2 + 1]) / g.dz;
        div[i] = dvxdx + dvzdz;
        exx[i] = dvxdx;
        ezz[i] = dvzdz;
        return;
    }
    i -= nc;
    if (i < nvtx) {
        const int k = i % Nx, l = i / Nx;
        const int c1 = k + l * Nx, c2 = k + l * (Nx + 1);
        const double dvxdz = (u[c1 + Nx] - u[c1]) / g.dz;
        const double dvzdx = (v[c2 + 1] - v[c2]) / g.dx;
        exz[i] = 0.5 * (dvxdz + dvzdx);
    }
}

// Momentum and continuity residuals from stresses
// (StokesAssemblyDecoupled.c:184-211, 688-724, 1128-1144), written in
// equation numbering, with the d scaling and interior RMS norms of
// StokesRoutines.c:890-932.  Partial sums of squares go to part[3*block].
__global__ void __launch_bounds__(256) k_residual(sgrid g, const double *__restrict__ sxx,
                                                  const double *__restrict__ szz,
                                                  const double *__restrict__ sxz,
                                                  const double *__restrict__ pc,
                                                  const double *__restrict__ div,
                                                  const double *__restrict__ rgx,
                                                  const double *__restrict__ rgz,
                                                  const double *__restrict__ dsc,
                                                  const int *__restrict__ eu,
                                                  const int *__restrict__ ev,
                                                  const int *__restrict__ ep,
                                                  double *__restrict__ F, double *__restrict__ part) {
    __shared__ double red[3][256];
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Ncx = Nx - 1, nu = nvx(g), nv = nvz(g), np = ncell(g);
    const double celvol = g.dx * g.dz;
    double sx = 0.0, sz = 0.0, sp_ = 0.0;
    if (i < nu) {
        const int k = i % Nx, l = i / Nx;
        const int e = eu[i];
        if (e >= 0 && l > 0 && l < g.Nz) {
            const int c2 = k - 1 + (l - 1) * Ncx;
            double f = (sxx[c2 + 1] - sxx[c2]) / g.dx;
            f -= (pc[c2 + 1] - pc[c2]) / g.dx;
            f += (sxz[i] - sxz[i - Nx]) / g.dz;
            f += rgx[i];
            f *= -celvol;
            if (dsc) f *= dsc[e];
            F[e] = f;
            if (l > 2 && l < g.Nz + 1 - 2) sx = f * f;
        }
    } else if ((i -= nu) < nv) {
        const int k = i % (Nx + 1), l = i / (Nx + 1);
        const int e = ev[i];
        if (e >= 0 && k > 0 && k < Nx) {
            const int c1 = k + l * Nx, c2 = k - 1 + (l - 1) * Ncx;
            double f = (szz[c2 + Ncx] - szz[c2]) / g.dz;
            f -= (pc[c2 + Ncx] - pc[c2]) / g.dz;
            f += (sxz[c1] - sxz[c1 - 1]) / g.dx;
            f += rgz[i];
            f *= -celvol;
            if (dsc) f *= dsc[e];
            F[e] = f;
            if (k > 2 && k < Nx + 1 - 2) sz = f * f;
        }
    } else if ((i -= nv) < np) {
        const int e = ep[i];
        if (e >= 0) {
            double f = celvol * div[i];
            sp_ = f * f;               // continuity norm before scaling
            if (dsc) f *= dsc[e];
            F[e] = f;
        }
    }
    red[0][threadIdx.x] = sx; red[1][threadIdx.x] = sz; red[2][threadIdx.x] = sp_;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int q = 0; q < 3; q++) part[3 * blockIdx.x + q] = red[q][0];
}

// Body force (EvaluateRHS, StokesRoutines.c:983-1100, no polar, no surface processes)
__global__ void k_body_force(sgrid g, double gx, double gz, const double *__restrict__ rho_n,
                             const double *__restrict__ rho_s, const int *__restrict__ eu,
                             const int *__restrict__ ev, double *__restrict__ rgx,
                             double *__restrict__ rgz) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Ncx = Nx - 1, nu = nvx(g), nv = nvz(g);
    if (i < nu) {
        const int l = i / Nx;
        rgx[i] = (eu[i] >= 0 && l > 0 && l < g.Nz) ? gx * (0.5 * (rho_s[i - Nx] + rho_s[i])) : 0.0;
        return;
    }
    i -= nu;
    if (i < nv) {
        const int k = i % (Nx + 1), l = i / (Nx + 1);
        const int c2 = k - 1 + (l - 1) * Ncx;
        rgz[i] = (ev[i] >= 0 && k > 0 && k < Nx) ? gz * (0.5 * (rho_n[c2] + rho_n[c2 + Ncx])) : 0.0;
    }
}

// Ghost velocities from the tags (ApplyBC, StokesRoutines.c:415-500) for the
// supported closed-box layout: free slip copies the inner value, no slip
// reflects it about the stored 2*value.
__global__ void k_apply_bc(sgrid g, const signed char *__restrict__ tu,
                           const signed char *__restrict__ tv, const double *__restrict__ bu,
                           const double *__restrict__ bv, double *__restrict__ u,
                           double *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int Nx = g.Nx, Nz = g.Nz, NxV = Nx + 1;
    if (i < Nx) {
        const int s = i, n = i + Nz * Nx;
        if (tu[s] == 13) u[s] = u[s + Nx];
        if (tu[n] == 13) u[n] = u[n - Nx];
        if (tu[s] == 11) u[s] = bu[s] - u[s + Nx];
        if (tu[n] == 11) u[n] = bu[n] - u[n - Nx];
    }
    if (i < Nz) {
        const int w = i * NxV, e = i * NxV + NxV - 1;
        if (tv[w] == 13) v[w] = v[w + 1];
        if (tv[e] == 13) v[e] = v[e - 1];
        if (tv[w] == 11) v[w] = bv[w] - v[w + 1];
        if (tv[e] == 11) v[e] = bv[e] - v[e - 1];
    }
}

template <typename T>
static int out_buf(mdg_ctx *c, T *user, size_t n, dbuf<T> &tmp, T **dev) {
    if (mdg_is_device_ptr(user)) { *dev = user; return MDG_OK; }
    HIPCHK(c, tmp.reserve(n));
    *dev = tmp.p;
    return MDG_OK;
}

extern "C" int mdg_apply_bc(mdg_ctx *c, double *u, double *v) {
    if (!c || !u || !v) return mdg_fail(c, MDG_ERR_ARG, "null velocity");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const sgrid g = c->sg;
    const bool dev = mdg_is_device_ptr(u) && mdg_is_device_ptr(v);
    double *du = u, *dv = v;
    if (!dev) {
        HIPCHK(c, c->tmp_a.reserve(nvx(g) + nvz(g)));
        du = c->tmp_a.p; dv = du + nvx(g);
        HIPCHK(c, hipMemcpyAsync(du, u, nvx(g) * 8, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dv, v, nvz(g) * 8, hipMemcpyHostToDevice, c->stream));
    }
    const int n = g.Nx > g.Nz ? g.Nx : g.Nz;
    hipLaunchKernelGGL(k_apply_bc, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, c->BCu_t.p,
                       c->BCv_t.p, c->BCu_v.p, c->BCv_v.p, du, dv);
    HIPCHK(c, hipGetLastError());
    if (!dev) {
        HIPCHK(c, hipMemcpyAsync(u, du, nvx(g) * 8, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(v, dv, nvz(g) * 8, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

// ExtractSolutions2 (StokesRoutines.c:536-576): u, v, p += alpha * the
// correction on every node that carries an unknown (the reference's tag test
// selects exactly the numbered nodes of a closed box), then ApplyBC.
__global__ void k_extract_add(int nu, int nv, int nc, const int *__restrict__ eu, const int *__restrict__ ev,
                              const int *__restrict__ ep, const double *__restrict__ x, double alpha,
                              double *__restrict__ u, double *__restrict__ v, double *__restrict__ p) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nu) {
        const int e = eu[i];
        if (e >= 0) u[i] = u[i] + alpha * x[e];
    } else if (i < nu + nv) {
        const int e = ev[i - nu];
        if (e >= 0) v[i - nu] = v[i - nu] + alpha * x[e];
    } else if (i < nu + nv + nc) {
        const int e = ep[i - nu - nv];
        if (e >= 0) p[i - nu - nv] = p[i - nu - nv] + alpha * x[e];
    }
}

extern "C" int mdg_extract_solution(mdg_ctx *c, const double *sol, double alpha, double *u, double *v, double *p) {
    if (!c || !sol || !u || !v || !p) return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const sgrid g = c->sg;
    const int nu = nvx(g), nv = nvz(g), nc = ncell(g);
    const double *dx;
    int rc;
    if ((rc = mdg_stage_in(c, sol, (size_t)c->neq, c->eqbuf, &dx))) return rc;
    double *user[3] = {u, v, p}, *d[3];
    const size_t sz[3] = {(size_t)nu, (size_t)nv, (size_t)nc};
    const bool dev = mdg_is_device_ptr(u) && mdg_is_device_ptr(v) && mdg_is_device_ptr(p);
    if (dev) {
        for (int q = 0; q < 3; q++) d[q] = user[q];
    } else {
        HIPCHK(c, c->tmp_a.reserve(sz[0] + sz[1] + sz[2]));
        d[0] = c->tmp_a.p; d[1] = d[0] + sz[0]; d[2] = d[1] + sz[1];
        for (int q = 0; q < 3; q++)
            HIPCHK(c, hipMemcpyAsync(d[q], user[q], sz[q] * 8, mdg_is_device_ptr(user[q]) ? hipMemcpyDeviceToDevice
                                                                                          : hipMemcpyHostToDevice,
                                     c->stream));
    }
    hipLaunchKernelGGL(k_extract_add, dim3(mdg_grid1d(nu + nv + nc, 256)), dim3(256), 0, c->stream, nu, nv, nc,
                       c->eqn_u.p, c->eqn_v.p, c->eqn_p.p, dx, alpha, d[0], d[1], d[2]);
    const int n = g.Nx > g.Nz ? g.Nx : g.Nz;
    hipLaunchKernelGGL(k_apply_bc, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, c->BCu_t.p,
                       c->BCv_t.p, c->BCu_v.p, c->BCv_v.p, d[0], d[1]);
    HIPCHK(c, hipGetLastError());
    if (!dev)
        for (int q = 0; q < 3; q++)
            HIPCHK(c, hipMemcpyAsync(user[q], d[q], sz[q] * 8, mdg_is_device_ptr(user[q]) ? hipMemcpyDeviceToDevice
                                                                                          : hipMemcpyDeviceToHost,
                                     c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_strain_rates(mdg_ctx *c, const double *u, const double *v, double *div,
                                double *exxd, double *ezzd, double *exz) {
    if (!c || !u || !v || !div || !exxd || !ezzd || !exz) return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const sgrid g = c->sg;
    const int nc = ncell(g), nvt = g.Nx * g.Nz;
    const double *du, *dv;
    int rc;
    if ((rc = mdg_stage_in(c, u, (size_t)nvx(g), c->tmp_a, &du)) || (rc = mdg_stage_in(c, v, (size_t)nvz(g), c->tmp_b, &dv)))
        return rc;
    double *o[4];
    dbuf<double> *tmps[4] = {&c->gx, &c->gy, &c->tmp_c, &c->tmp_d};
    double *user[4] = {div, exxd, ezzd, exz};
    const size_t sz[4] = {(size_t)nc, (size_t)nc, (size_t)nc, (size_t)nvt};
    for (int q = 0; q < 4; q++)
        if ((rc = out_buf(c, user[q], sz[q], *tmps[q], &o[q]))) return rc;
    hipLaunchKernelGGL(k_strain_rates, dim3(mdg_grid1d(nc + nvt, 256)), dim3(256), 0, c->stream, g, du,
                       dv, o[0], o[1], o[2], o[3]);
    HIPCHK(c, hipGetLastError());
    for (int q = 0; q < 4; q++)
        if (o[q] != user[q])
            HIPCHK(c, hipMemcpyAsync(user[q], o[q], sz[q] * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_eval_rhs(mdg_ctx *c, double gx, double gz, const double *rho_n,
                            const double *rho_s, double *roger_x, double *roger_z) {
    if (!c || !rho_n || !rho_s || !roger_x || !roger_z) return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const sgrid g = c->sg;
    const double *rn, *rs;
    int rc;
    if ((rc = mdg_stage_in(c, rho_n, (size_t)ncell(g), c->tmp_a, &rn)) ||
        (rc = mdg_stage_in(c, rho_s, (size_t)g.Nx * g.Nz, c->tmp_b, &rs)))
        return rc;
    double *ox, *oz;
    if ((rc = out_buf(c, roger_x, (size_t)nvx(g), c->tmp_c, &ox)) || (rc = out_buf(c, roger_z, (size_t)nvz(g), c->tmp_d, &oz)))
        return rc;
    hipLaunchKernelGGL(k_body_force, dim3(mdg_grid1d(nvx(g) + nvz(g), 256)), dim3(256), 0, c->stream, g,
                       gx, gz, rn, rs, c->eqn_u.p, c->eqn_v.p, ox, oz);
    HIPCHK(c, hipGetLastError());
    if (ox != roger_x) HIPCHK(c, hipMemcpyAsync(roger_x, ox, nvx(g) * 8, hipMemcpyDeviceToHost, c->stream));
    if (oz != roger_z) HIPCHK(c, hipMemcpyAsync(roger_z, oz, nvz(g) * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

extern "C" int mdg_stokes_residual(mdg_ctx *c, const double *sxxd, const double *szzd,
                                   const double *sxz, const double *p_corr, const double *div,
                                   const double *roger_x, const double *roger_z, const double *d,
                                   double *F, double norms[3]) {
    if (!c || !sxxd || !szzd || !sxz || !p_corr || !div || !roger_x || !roger_z || !F || !norms)
        return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    const sgrid g = c->sg;
    const size_t nc = ncell(g), nvt = (size_t)g.Nx * g.Nz;
    // Stage all inputs in one device buffer when they live on the host.
    const double *in[7] = {sxxd, szzd, sxz, p_corr, div, roger_x, roger_z};
    const size_t sz[7] = {nc, nc, nvt, nc, nc, (size_t)nvx(g), (size_t)nvz(g)};
    size_t tot = 0;
    for (int q = 0; q < 7; q++) tot += sz[q];
    HIPCHK(c, c->gx.reserve(tot + (size_t)c->neq));
    const double *dp[7];
    size_t off = 0;
    for (int q = 0; q < 7; q++) {
        if (mdg_is_device_ptr(in[q])) dp[q] = in[q];
        else {
            HIPCHK(c, hipMemcpyAsync(c->gx.p + off, in[q], sz[q] * 8, hipMemcpyHostToDevice, c->stream));
            dp[q] = c->gx.p + off;
        }
        off += sz[q];
    }
    const double *dd = nullptr;
    int rc;
    if (d && (rc = mdg_stage_in(c, d, (size_t)c->neq, c->tmp_a, &dd))) return rc;
    double *Fd;
    if ((rc = out_buf(c, F, (size_t)c->neq, c->eqbuf, &Fd))) return rc;
    const int n = nvx(g) + nvz(g) + (int)nc;
    const int nb = mdg_grid1d(n, 256);
    HIPCHK(c, c->part.reserve(3 * (size_t)nb));
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, c->stream, g, dp[0], dp[1], dp[2], dp[3],
                       dp[4], dp[5], dp[6], dd, c->eqn_u.p, c->eqn_v.p, c->eqn_p.p, Fd, c->part.p);
    HIPCHK(c, hipGetLastError());
    std::vector<double> part(3 * (size_t)nb);
    HIPCHK(c, hipMemcpyAsync(part.data(), c->part.p, part.size() * 8, hipMemcpyDeviceToHost, c->stream));
    if (Fd != F) HIPCHK(c, hipMemcpyAsync(F, Fd, (size_t)c->neq * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double s[3] = {0, 0, 0};
    for (int b = 0; b < nb; b++)
        for (int q = 0; q < 3; q++) s[q] += part[3 * b + q];
    norms[0] = sqrt(s[0] / c->ndof_norm[0]);
    norms[1] = sqrt(s[1] / c->ndof_norm[1]);
    norms[2] = sqrt(s[2] / (double)c->neq_cont);
    if (std::isnan(norms[0]) || std::isnan(norms[1]) || std::isnan(norms[2]))
        return mdg_fail(c, MDG_ERR_NAN, "NaN residual");
    return MDG_OK;
}

// ------------------------------------------------ allocation-free residual
// Device scratch of mdg_stokes_residual_static: the residual norms' partial
// sums (3 per block) and their totals.  One copy per GPU, held until the
// process ends; every call on that GPU shares it, so calls that overlap on
// one GPU race.  Host code never uses the array's name as an address: the
// address comes from hipGetSymbolAddress, once per device, at context set-up.
static __device__ __align__(256) double scratch[MDG_STATIC_SCRATCH_DOUBLES];
static double *g_scratch_dev[MDG_MAX_DEVICES];
static std::mutex g_scratch_mu;

int mdg_static_scratch_init(mdg_ctx *c) {
    int dev = 0;
    HIPCHK(c, hipGetDevice(&dev));
    if (dev < 0 || dev >= MDG_MAX_DEVICES) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        HIPCHK(c, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return mdg_fail(c, MDG_ERR_HIP, "no device address for the static scratch");
        g_scratch_dev[dev] = static_cast<double *>(p);
    }
    c->static_scratch = g_scratch_dev[dev];
    return MDG_OK;
}

// totals of the 3 x nb partial sums, each summed in block order (the order
// of mdg_stokes_residual's host loop, so both give the same bits)
__global__ void k_sum3_ordered(int nb, const double *__restrict__ part, double *__restrict__ tot) {
    const int q = threadIdx.x;
    if (q >= 3) return;
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += part[3 * b + q];
    tot[q] = s;
}

extern "C" int mdg_stokes_residual_static(mdg_ctx *c, const double *sxxd, const double *szzd,
                                          const double *sxz, const double *p_corr, const double *div,
                                          const double *roger_x, const double *roger_z,
                                          const double *d, double *F, double norms[3]) {
    if (!c || !sxxd || !szzd || !sxz || !p_corr || !div || !roger_x || !roger_z || !F || !norms)
        return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (!c->static_scratch) return mdg_fail(c, MDG_ERR_STATE, "no static scratch on this device");
    const double *in[9] = {sxxd, szzd, sxz, p_corr, div, roger_x, roger_z, d, F};
    for (int q = 0; q < 9; q++)
        if (in[q] && !mdg_is_device_ptr(in[q]))
            return mdg_fail(c, MDG_ERR_ARG, "mdg_stokes_residual_static takes device arrays only (argument %d)", q + 2);
    const sgrid g = c->sg;
    const int n = nvx(g) + nvz(g) + ncell(g);
    const int nb = mdg_grid1d(n, 256);
    const long need = 3L * nb + 3;
    if (need > MDG_STATIC_SCRATCH_DOUBLES)
        return mdg_fail(c, MDG_ERR_ARG, "grid too large for the static scratch: %ld partial sums, at most %d", need,
                        MDG_STATIC_SCRATCH_DOUBLES);
    double *part = c->static_scratch, *tot = part + 3L * nb;
    hipLaunchKernelGGL(k_residual, dim3(nb), dim3(256), 0, c->stream, g, sxxd, szzd, sxz, p_corr, div, roger_x,
                       roger_z, d, c->eqn_u.p, c->eqn_v.p, c->eqn_p.p, F, part);
    hipLaunchKernelGGL(k_sum3_ordered, dim3(1), dim3(64), 0, c->stream, nb, (const double *)part, tot);
    HIPCHK(c, hipGetLastError());
    double s[3];
    HIPCHK(c, hipMemcpyAsync(s, tot, sizeof(s), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    norms[0] = sqrt(s[0] / c->ndof_norm[0]);
    norms[1] = sqrt(s[1] / c->ndof_norm[1]);
    norms[2] = sqrt(s[2] / (double)c->neq_cont);
    if (std::isnan(norms[0]) || std::isnan(norms[1]) || std::isnan(norms[2]))
        return mdg_fail(c, MDG_ERR_NAN, "NaN residual");
    return MDG_OK;
}

// ExtractDiagonalScale (StokesRoutines.c:1260-1300): d = the diagonal of the
// assembled Picard block A on momentum rows, and 1 on continuity rows (the D
// block is empty for incompressible cells, so its d keeps SAlloc's 1.0).
__global__ void k_stokes_diag(sgrid g, const double *__restrict__ en, const double *__restrict__ es,
                              const int *__restrict__ eu, const int *__restrict__ ev,
                              const int *__restrict__ ep, double *__restrict__ d) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = nvx(g), nw = nvz(g), nc = ncell(g);
    if (i < nu) {
        const int e = eu[i];
        if (e >= 0) d[e] = diag_vx<double>(g, i % g.Nx, i / g.Nx, en, es, 0.0);
    } else if (i < nu + nw) {
        const int j = i - nu, e = ev[j];
        if (e >= 0) d[e] = diag_vz<double>(g, j % (g.Nx + 1), j / (g.Nx + 1), en, es, 0.0);
    } else if (i < nu + nw + nc) {
        const int e = ep[i - nu - nw];
        if (e >= 0) d[e] = 1.0;
    }
}

extern "C" int mdg_stokes_diag(mdg_ctx *c, const double *D11_n, const double *D22_n,
                               const double *D33_s, double *d) {
    if (!c || !d) return mdg_fail(c, MDG_ERR_ARG, "null vector");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    int rc;
    if ((rc = mdg_stage_coeffs(c, D11_n, D22_n, D33_s)) || (rc = mdg_reserve_grid_vectors(c))) return rc;
    const sgrid g = c->sg;
    double *dd = mdg_is_device_ptr(d) ? d : c->eqbuf.p;
    const int n = nvx(g) + nvz(g) + ncell(g);
    hipLaunchKernelGGL(k_stokes_diag, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, c->eta_n.p,
                       c->eta_s.p, c->eqn_u.p, c->eqn_v.p, c->eqn_p.p, dd);
    HIPCHK(c, hipGetLastError());
    if (dd != d) HIPCHK(c, hipMemcpyAsync(d, dd, (size_t)c->neq * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}
