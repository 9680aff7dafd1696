// Context lifetime, grid coordinates, boundary tags, error reporting.
#include "stokes_internal.h"

int mdg_fail(mdg_ctx *c, int code, const char *fmt, ...) {
    if (c) {
        char buf[512];
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        c->err = buf;
    }
    return code;
}

bool mdg_is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t a;
    hipError_t e = hipPointerGetAttributes(&a, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();  // unregistered host memory reports an error
        return false;
    }
    return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Running sums as in SetGridCoordinates (GridRoutines.c:246-330).
static void running(std::vector<double> &v, double x0, double h, int n) {
    v.resize(n);
    v[0] = x0;
    for (int i = 1; i < n; i++) v[i] = v[i - 1] + h;
}

template <typename T>
static int upload(mdg_ctx *c, dbuf<T> &d, const std::vector<T> &h) {
    HIPCHK(c, d.reserve(h.size()));
    HIPCHK(c, hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return MDG_OK;
}

extern "C" int mdg_create(const mdg_grid_desc *d, mdg_ctx **out) {
    if (!out) return MDG_ERR_ARG;
    *out = nullptr;
    if (!d || d->Nx < 3 || d->Nz < 3 || !(d->xmax > d->xmin) || !(d->zmax > d->zmin))
        return MDG_ERR_ARG;
    if (d->periodic_x) return MDG_ERR_UNSUPPORTED;
    mdg_ctx *c = new mdg_ctx();
    c->Nx = d->Nx; c->Nz = d->Nz; c->Ncx = d->Nx - 1; c->Ncz = d->Nz - 1;
    c->xmin = d->xmin; c->xmax = d->xmax; c->zmin = d->zmin; c->zmax = d->zmax;
    c->dx = (d->xmax - d->xmin) / (d->Nx - 1);
    c->dz = (d->zmax - d->zmin) / (d->Nz - 1);
    running(c->h_xg, d->xmin, c->dx, c->Nx);
    running(c->h_zg, d->zmin, c->dz, c->Nz);
    running(c->h_xc, d->xmin + c->dx / 2.0, c->dx, c->Nx - 1);
    running(c->h_zc, d->zmin + c->dz / 2.0, c->dz, c->Nz - 1);
    running(c->h_xvz, d->xmin - c->dx / 2.0, c->dx, c->Nx + 1);
    running(c->h_zvx, d->zmin - c->dz / 2.0, c->dz, c->Nz + 1);
    int rc = MDG_OK;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return MDG_ERR_HIP;
    }
    if ((rc = upload(c, c->xg, c->h_xg)) || (rc = upload(c, c->zg, c->h_zg)) ||
        (rc = upload(c, c->xc, c->h_xc)) || (rc = upload(c, c->zc, c->h_zc)) ||
        (rc = upload(c, c->xvz, c->h_xvz)) || (rc = upload(c, c->zvx, c->h_zvx)) ||
        (rc = (c->flag_dev.reserve(16) == hipSuccess) ? MDG_OK : MDG_ERR_HIP) ||
        (rc = mdg_static_scratch_init(c))) {
        mdg_destroy(c);
        return rc;
    }
    *out = c;
    return MDG_OK;
}

extern "C" void mdg_destroy(mdg_ctx *c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    c->xg.release(); c->zg.release(); c->xc.release(); c->zc.release();
    c->xvz.release(); c->zvx.release();
    c->BCu_t.release(); c->BCv_t.release(); c->BCp_t.release(); c->BCg_t.release();
    c->BCu_v.release(); c->BCv_v.release();
    c->mx.release(); c->mz.release(); c->mphase.release(); c->perm.release();
    c->qoff.release(); c->keys.release(); c->keys2.release(); c->idx.release();
    for (auto &b : c->mw) b.release();
    c->sort_tmp.release(); c->field_tmp.release(); c->node_tmp.release();
    c->table_tmp.release(); c->ntype_tmp.release(); c->flag_dev.release();
    c->eqn_u.release(); c->eqn_v.release(); c->eqn_p.release(); c->eta_n.release();
    c->eta_s.release(); c->tmp_a.release(); c->tmp_b.release(); c->tmp_c.release();
    c->tmp_d.release(); c->gx.release(); c->gy.release(); c->eqbuf.release(); c->part.release();
    mdg_solver_release(c);
    mdg_rheo_release(c);
    mdg_pcnet_release(c);
    mdg_newton_release(c);
    mdg_advect_release(c);
    mdg_g2p_release(c);
    mdg_thermal_release(c);
    mdg_reseed_release(c);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char *mdg_last_error(const mdg_ctx *c) { return c ? c->err.c_str() : "null context"; }

extern "C" int mdg_grid_coords(const mdg_ctx *c, double *xg, double *zg, double *xc, double *zc,
                               double *xvz, double *zvx) {
    if (!c) return MDG_ERR_ARG;
    auto cp = [](double *dst, const std::vector<double> &s) {
        if (dst) for (size_t i = 0; i < s.size(); i++) dst[i] = s[i];
    };
    cp(xg, c->h_xg); cp(zg, c->h_zg); cp(xc, c->h_xc); cp(zc, c->h_zc);
    cp(xvz, c->h_xvz); cp(zvx, c->h_zvx);
    return MDG_OK;
}

template <typename T>
static int copy_any(mdg_ctx *c, dbuf<T> &d, const T *src, size_t n) {
    if (!src) return mdg_fail(c, MDG_ERR_ARG, "null boundary array");
    HIPCHK(c, d.reserve(n));
    HIPCHK(c, hipMemcpyAsync(d.p, src, n * sizeof(T), hipMemcpyDefault, c->stream));
    return MDG_OK;
}

extern "C" int mdg_set_bc(mdg_ctx *c, const signed char *BCu_t, const double *BCu_v,
                          const signed char *BCv_t, const double *BCv_v,
                          const signed char *BCp_t, const signed char *BCg_t) {
    if (!c) return MDG_ERR_ARG;
    const size_t nu = (size_t)c->Nx * (c->Nz + 1), nv = (size_t)(c->Nx + 1) * c->Nz;
    const size_t np = (size_t)c->Ncx * c->Ncz, ng = (size_t)c->Nx * c->Nz;
    // The implemented stencils cover tags -1, 0, 11, 13 (no air, stress or
    // periodic rows); reject anything else up front.
    auto check = [&](const signed char *t, size_t n, const char *name) -> int {
        std::vector<signed char> h(n);
        if (hipMemcpy(h.data(), t, n, hipMemcpyDefault) != hipSuccess)
            return mdg_fail(c, MDG_ERR_HIP, "reading %s", name);
        for (size_t i = 0; i < n; i++)
            if (!(h[i] == -1 || h[i] == 0 || h[i] == 11 || h[i] == 13))
                return mdg_fail(c, MDG_ERR_UNSUPPORTED, "%s[%zu] = %d: only tags -1, 0, 11, 13 "
                                "are implemented", name, i, (int)h[i]);
        return MDG_OK;
    };
    int rc;
    if ((rc = check(BCu_t, nu, "BCu_t")) || (rc = check(BCv_t, nv, "BCv_t"))) return rc;
    if ((rc = copy_any(c, c->BCu_t, BCu_t, nu)) || (rc = copy_any(c, c->BCu_v, BCu_v, nu)) ||
        (rc = copy_any(c, c->BCv_t, BCv_t, nv)) || (rc = copy_any(c, c->BCv_v, BCv_v, nv)) ||
        (rc = copy_any(c, c->BCp_t, BCp_t, np)) || (rc = copy_any(c, c->BCg_t, BCg_t, ng)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::vector<signed char> hu(nu), hv(nv), hp(np), hg(ng);
    HIPCHK(c, hipMemcpy(hu.data(), c->BCu_t.p, nu, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hv.data(), c->BCv_t.p, nv, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hp.data(), c->BCp_t.p, np, hipMemcpyDeviceToHost));
    HIPCHK(c, hipMemcpy(hg.data(), c->BCg_t.p, ng, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < np; i++)
        if (hp[i] != -1) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "BCp_t must be -1 (no air cells)");
    for (size_t i = 0; i < ng; i++)
        if (hg[i] == 30) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "BCg_t 30 (air) not implemented");
    // The matrix-free stencils assume MDOODZ's closed-box layout: interior
    // -1, Dirichlet normal velocity on the walls (tag 0), and one tangential
    // condition per side on the ghost rows/columns (13 free slip or 11 no slip).
    const int Nx = c->Nx, Nz = c->Nz;
    int side[4] = {0, 0, 0, 0};   // S, N, W, E: 1 = tag 13, 2 = tag 11
    auto kind = [](signed char t) { return t == 13 ? 1 : (t == 11 ? 2 : 0); };
    for (int l = 0; l < Nz + 1; l++)
        for (int k = 0; k < Nx; k++) {
            const signed char t = hu[k + l * Nx];
            const bool wall = (k == 0 || k == Nx - 1), ghost = (l == 0 || l == Nz);
            if (ghost && !wall) {
                const int s = (l == 0) ? 0 : 1, kd = kind(t);
                if (!kd || (side[s] && side[s] != kd))
                    return mdg_fail(c, MDG_ERR_UNSUPPORTED, "Vx ghost row %d: tag %d (need one of 13/11 per side)", l, (int)t);
                side[s] = kd;
            } else if (wall && !ghost) {
                if (t != 0) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "Vx wall node (%d,%d) tag %d (need 0)", k, l, (int)t);
            } else if (!wall && !ghost && t != -1) {
                return mdg_fail(c, MDG_ERR_UNSUPPORTED, "interior Vx node (%d,%d) tag %d", k, l, (int)t);
            }
        }
    for (int l = 0; l < Nz; l++)
        for (int k = 0; k < Nx + 1; k++) {
            const signed char t = hv[k + l * (Nx + 1)];
            const bool wall = (l == 0 || l == Nz - 1), ghost = (k == 0 || k == Nx);
            if (ghost && !wall) {
                const int s = (k == 0) ? 2 : 3, kd = kind(t);
                if (!kd || (side[s] && side[s] != kd))
                    return mdg_fail(c, MDG_ERR_UNSUPPORTED, "Vz ghost column %d: tag %d (need one of 13/11 per side)", k, (int)t);
                side[s] = kd;
            } else if (wall && !ghost) {
                if (t != 0) return mdg_fail(c, MDG_ERR_UNSUPPORTED, "Vz wall node (%d,%d) tag %d (need 0)", k, l, (int)t);
            } else if (!wall && !ghost && t != -1) {
                return mdg_fail(c, MDG_ERR_UNSUPPORTED, "interior Vz node (%d,%d) tag %d", k, l, (int)t);
            }
        }
    c->sg = sgrid{Nx, Nz, c->dx, c->dz, side[0], side[1], side[2], side[3]};
    int rc2 = mdg_build_numbering(c, hu, hv, hp);
    if (rc2) return rc2;
    c->have_bc = true;
    return MDG_OK;
}

int mdg_node_set(mdg_ctx *c, int centroid, node_set *ns) {
    switch (centroid) {
    case 1:  *ns = {c->Ncx, c->Ncz, c->xc.p, c->zc.p, c->h_xc[0], c->h_zc[0]}; break;
    case 0:  *ns = {c->Nx, c->Nz, c->xg.p, c->zg.p, c->h_xg[0], c->h_zg[0]}; break;
    case -1: *ns = {c->Nx, c->Nz + 1, c->xg.p, c->zvx.p, c->h_xg[0], c->h_zvx[0]}; break;
    case -2: *ns = {c->Nx + 1, c->Nz, c->xvz.p, c->zg.p, c->h_xvz[0], c->h_zg[0]}; break;
    default: return mdg_fail(c, MDG_ERR_ARG, "centroid %d not in {1,0,-1,-2}", centroid);
    }
    return MDG_OK;
}

extern "C" int mdg_last_kernel_ms(const mdg_ctx *c, double *ms) {
    if (!c || !ms) return MDG_ERR_ARG;
    *ms = c->last_ms;
    return MDG_OK;
}

extern "C" int mdg_last_weights_ms(const mdg_ctx *c, double *ms) {
    if (!c || !ms) return MDG_ERR_ARG;
    *ms = c->weights_ms;
    return MDG_OK;
}

extern "C" int mdg_dev_alloc(mdg_ctx *c, unsigned long long bytes, void **p) {
    if (!c || !p) return MDG_ERR_ARG;
    *p = nullptr;
    HIPCHK(c, hipMalloc(p, bytes ? bytes : 1));
    return MDG_OK;
}

extern "C" int mdg_dev_free(mdg_ctx *c, void *p) {
    if (!c) return MDG_ERR_ARG;
    if (p) HIPCHK(c, hipFree(p));
    return MDG_OK;
}

extern "C" int mdg_memcpy(mdg_ctx *c, void *dst, const void *src, unsigned long long bytes) {
    if (!c || (bytes && (!dst || !src))) return MDG_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

__global__ void k_dev_scale(long n, double a, const double *__restrict__ x, double *__restrict__ y) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i];
}

extern "C" int mdg_dev_scale(mdg_ctx *c, long long n, double a, const double *x, double *y) {
    if (!c || n < 0 || (n && (!x || !y))) return MDG_ERR_ARG;
    if (n == 0) return MDG_OK;
    if (!mdg_is_device_ptr(x) || !mdg_is_device_ptr(y)) return mdg_fail(c, MDG_ERR_ARG, "device arrays expected");
    hipLaunchKernelGGL(k_dev_scale, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, (long)n, a, x, y);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}
