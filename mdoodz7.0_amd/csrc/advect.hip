// Marker advection: V2P (ParticleRoutines.c:63-256) and the fourth-order
// Runge-Kutta step of RogerGuntherII (AdvectionRoutines.c:347-410) with
// isoutPart (AdvectionRoutines.c:320-341, non-periodic box).
//
// One thread per marker: the four V2P stages gather 2 x 4 velocities each
// from the Vx / Vz grids (read through L2; a sorted marker set keeps the
// gathers of a wavefront within a few grid rows) and the marker streams
// x, z, phase in and x, z, phase, vx, vz out: 44 B of HBM traffic per
// marker. The arithmetic is the reference's, in the same order, built
// without contraction, so positions and velocities match the CPU
// restatement bit for bit.
#include "mdg_internal.h"

namespace {

struct v2p_grid {
    int Nx, Nz;                       // vertices
    double dx, dz;
    const double *xg, *zg, *xvz, *zvx;
    const signed char *tu, *tv;       // BCu / BCv tags (30: air)
    double xmin, xmax, zmin, zmax;    // box of isoutPart
};

__device__ inline int v2p_cell(double d, double h, int n) {
    int j = (int)ceil(d / h) - 1;
    if (j < 0) j = 0;
    if (j > n - 2) j = n - 2;
    return j;
}
// 0.5 pow(t - 0.5, 2): the square written as one product, the correctly
// rounded value the C library's pow returns for an exponent of 2
__device__ inline double v2p_corr(double t) {
    const double y = t - 0.5;
    return 0.5 * (y * y);
}

// V2P of one active marker (ParticleRoutines.c:76-252)
__device__ void v2p_one(const v2p_grid &g, const double *__restrict__ u, const double *__restrict__ v,
                        int conserv, double x, double z, double *vxm, double *vzm) {
    const double dx = g.dx, dz = g.dz;
    {
        const signed char *t = g.tu;
        const int Nx = g.Nx, Nz = g.Nz + 1;
        const int j = v2p_cell(x - g.xg[0], dx, Nx), i = v2p_cell(z - g.zvx[0], dz, Nz);
        const double dxm = fabs(x - g.xg[j]), dzm = fabs(z - g.zvx[i]);
        const int iSW = j + i * Nx, iSE = iSW + 1, iNW = j + (i + 1) * Nx, iNE = iNW + 1;
        double VNW = u[iNW], VNE = u[iNE], VSW = u[iSW], VSE = u[iSE];
        const bool aSW = t[iSW] == 30, aSE = t[iSE] == 30, aNW = t[iNW] == 30, aNE = t[iNE] == 30;
        if (aSW) VSW = u[iNW];
        if (aSE) VSE = u[iNE];
        if (aNW) VNW = u[iSW];
        if (aNE) VNE = u[iSE];
        if (aNE && aSE) { VNE = VSW; VSE = VSW; }
        if (aNW && aSW) { VNW = VSE; VSW = VSE; }
        double v13 = VSW * (1.0 - dxm / dx) + VSE * dxm / dx;
        double v24 = VNW * (1.0 - dxm / dx) + VNE * dxm / dx;
        if (conserv == 1) {
            if (dxm / dx >= 0.5) {
                if (j < Nx - 2) {
                    const int iSEE = iSE + 1, iNEE = iNE + 1;
                    double VSEE = u[iSEE], VNEE = u[iNEE];
                    if (t[iSEE] == 30) VSEE = u[iSW];
                    if (t[iNEE] == 30) VNEE = u[iNW];
                    v13 = v13 + v2p_corr(dxm / dx) * (VSW - 2.0 * VSE + VSEE);
                    v24 = v24 + v2p_corr(dxm / dx) * (VNW - 2.0 * VNE + VNEE);
                }
            } else if (j > 0) {
                const int iSWW = iSW - 1, iNWW = iNW - 1;
                double VSWW = u[iSWW], VNWW = u[iNWW];
                if (t[iSWW] == 30) VSWW = u[iSE];
                if (t[iNWW] == 30) VNWW = u[iNE];
                v13 = v13 + v2p_corr(dxm / dx) * (VSWW - 2.0 * VSW + VSE);
                v24 = v24 + v2p_corr(dxm / dx) * (VNWW - 2.0 * VNW + VNE);
            }
        }
        *vxm = (1.0 - dzm / dz) * v13 + (dzm / dz) * v24;
    }
    {
        const signed char *t = g.tv;
        const int Nx = g.Nx + 1, Nz = g.Nz;
        const int j = v2p_cell(x - g.xvz[0], dx, Nx), i = v2p_cell(z - g.zg[0], dz, Nz);
        const double dxm = fabs(x - g.xvz[j]), dzm = fabs(z - g.zg[i]);
        const int iSW = j + i * Nx, iSE = iSW + 1, iNW = j + (i + 1) * Nx, iNE = iNW + 1;
        double VSE = v[iSE], VSW = v[iSW], VNE = v[iNE], VNW = v[iNW];
        const bool aSW = t[iSW] == 30, aSE = t[iSE] == 30, aNW = t[iNW] == 30, aNE = t[iNE] == 30;
        // the reference replaces an air SW value by NE (ParticleRoutines.c:207)
        if (aSW) VSW = v[iNE];
        if (aSE) VSE = v[iNE];
        if (aNW) VNW = v[iSW];
        if (aNE) VNE = v[iSE];
        if (aNE && aSE) { VNE = VSW; VSE = VSW; }
        if (aNW && aSW) { VNW = VSE; VSW = VSE; }
        double v12 = VSW * (1.0 - dzm / dz) + VNW * dzm / dz;
        double v34 = VSE * (1.0 - dzm / dz) + VNE * dzm / dz;
        if (conserv == 1) {
            if (dzm / dz >= 0.5) {
                if (i < Nz - 2) {
                    const int iNNW = j + (i + 2) * Nx, iNNE = iNNW + 1;
                    double VNNW = v[iNNW], VNNE = v[iNNE];
                    if (t[iNNW] == 30) VNNW = v[iSW];
                    if (t[iNNE] == 30) VNNE = v[iSE];
                    v12 = v12 + v2p_corr(dzm / dz) * (VSW - 2.0 * VNW + VNNW);
                    v34 = v34 + v2p_corr(dzm / dz) * (VSE - 2.0 * VNE + VNNE);
                }
            } else if (i > 0) {
                const int iSSW = j + (i - 1) * Nx, iSSE = iSSW + 1;
                double VSSW = v[iSSW], VSSE = v[iSSE];
                if (t[iSSW] == 30) VSSW = v[iNW];
                if (t[iSSE] == 30) VSSE = v[iNE];
                v12 = v12 + v2p_corr(dzm / dz) * (VSSW - 2.0 * VSW + VNW);
                v34 = v34 + v2p_corr(dzm / dz) * (VSSE - 2.0 * VSE + VNE);
            }
        }
        *vzm = (1.0 - dxm / dx) * v12 + (dxm / dx) * v34;
    }
}

__device__ inline void isout(const v2p_grid &g, double x, double z, int *ph) {
    if (x < g.xmin || x > g.xmax || z < g.zmin || z > g.zmax) *ph = -1;
}

__global__ void k_v2p(v2p_grid g, const double *__restrict__ u, const double *__restrict__ v, int conserv, int n,
                      const double *__restrict__ x, const double *__restrict__ z, const int *__restrict__ phase,
                      double *__restrict__ vx, double *__restrict__ vz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double a = 0.0, b = 0.0;
    if (phase[k] != -1) v2p_one(g, u, v, conserv, x[k], z[k], &a, &b);
    vx[k] = a;
    vz[k] = b;
}

// RogerGuntherII's loop body (AdvectionRoutines.c:372-405)
__global__ void k_advect_rk4(v2p_grid g, const double *__restrict__ u, const double *__restrict__ v, double dt,
                             int conserv, int n, double *__restrict__ x, double *__restrict__ z,
                             int *__restrict__ phase, double *__restrict__ vxo, double *__restrict__ vzo) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double xA = x[k], zA = z[k];
    double xk = xA, zk = zA;
    int ph = phase[k];
    double VxA = 0.0, VzA = 0.0, VxB = 0.0, VzB = 0.0, VxC = 0.0, VzC = 0.0, VxD = 0.0, VzD = 0.0;
    if (ph != -1) v2p_one(g, u, v, conserv, xk, zk, &VxA, &VzA);
    if (ph != -1) {
        xk = xA + 0.5 * dt * VxA;
        zk = zA + 0.5 * dt * VzA;
    }
    isout(g, xk, zk, &ph);
    if (ph != -1) v2p_one(g, u, v, conserv, xk, zk, &VxB, &VzB);
    if (ph != -1) {
        xk = xA + 0.5 * dt * VxB;
        zk = zA + 0.5 * dt * VzB;
    }
    isout(g, xk, zk, &ph);
    if (ph != -1) v2p_one(g, u, v, conserv, xk, zk, &VxC, &VzC);
    if (ph != -1) {
        xk = xA + 1.0 * dt * VxC;
        zk = zA + 1.0 * dt * VzC;
    }
    isout(g, xk, zk, &ph);
    if (ph != -1) v2p_one(g, u, v, conserv, xk, zk, &VxD, &VzD);
    VxA = (1.0 / 6.0) * (VxA + 2.0 * VxB + 2.0 * VxC + VxD);
    VzA = (1.0 / 6.0) * (VzA + 2.0 * VzB + 2.0 * VzC + VzD);
    if (ph != -1) {
        xk = xA + dt * VxA;
        zk = zA + dt * VzA;
        if (vxo) vxo[k] = VxA;
        if (vzo) vzo[k] = VzA;
    }
    isout(g, xk, zk, &ph);
    x[k] = xk;
    z[k] = zk;
    phase[k] = ph;
}

v2p_grid make_grid(const mdg_ctx *c) {
    return v2p_grid{c->Nx, c->Nz, c->dx, c->dz, c->xg.p, c->zg.p, c->xvz.p, c->zvx.p,
                    c->BCu_t.p, c->BCv_t.p, c->xmin, c->xmax, c->zmin, c->zmax};
}

// A caller's array as a device pointer: itself, or a device copy (copied in
// when `in`); copy_back returns a device copy's contents to the caller.
template <typename T>
int dev_view(mdg_ctx *c, T *user, size_t n, dbuf<T> &tmp, bool in, T **d) {
    if (!user) { *d = nullptr; return MDG_OK; }
    if (mdg_is_device_ptr(user)) { *d = user; return MDG_OK; }
    HIPCHK(c, tmp.reserve(n));
    if (in) HIPCHK(c, hipMemcpyAsync(tmp.p, user, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    *d = tmp.p;
    return MDG_OK;
}
template <typename T>
int copy_back(mdg_ctx *c, T *user, const T *d, size_t n) {
    if (user && d && d != user) HIPCHK(c, hipMemcpyAsync(user, d, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    return MDG_OK;
}

}  // namespace

// staging buffers for host callers, one set per context (freed with it)
struct advect_scratch {
    dbuf<double> u, v, x, z, vx, vz;
    dbuf<int> ph;
};

static advect_scratch &scratch_of(mdg_ctx *c) {
    if (!c->advect) c->advect = new advect_scratch();
    return *c->advect;
}

void mdg_advect_release(mdg_ctx *c) {
    if (!c || !c->advect) return;
    advect_scratch *s = c->advect;
    s->u.release(); s->v.release(); s->x.release(); s->z.release(); s->vx.release(); s->vz.release(); s->ph.release();
    delete s;
    c->advect = nullptr;
}

extern "C" int mdg_v2p(mdg_ctx *c, const double *u, const double *v, int conserv_interp, int n, const double *x,
                       const double *z, const int *phase, double *vx, double *vz) {
    if (!c || !u || !v || !x || !z || !phase || !vx || !vz || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (n == 0) return MDG_OK;
    advect_scratch &s = scratch_of(c);
    const size_t nu = (size_t)c->Nx * (c->Nz + 1), nv = (size_t)(c->Nx + 1) * c->Nz;
    const double *du, *dv, *dx_, *dz_;
    const int *dph;
    double *dvx, *dvz;
    int rc;
    if ((rc = mdg_stage_in(c, u, nu, s.u, &du)) || (rc = mdg_stage_in(c, v, nv, s.v, &dv)) ||
        (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) || (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) ||
        (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)) || (rc = dev_view(c, vx, (size_t)n, s.vx, false, &dvx)) ||
        (rc = dev_view(c, vz, (size_t)n, s.vz, false, &dvz)))
        return rc;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_v2p, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, make_grid(c), du, dv, conserv_interp, n,
                       dx_, dz_, dph, dvx, dvz);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if ((rc = copy_back(c, vx, (const double *)dvx, (size_t)n)) || (rc = copy_back(c, vz, (const double *)dvz, (size_t)n)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return MDG_OK;
}

extern "C" int mdg_advect_rk4(mdg_ctx *c, const double *u, const double *v, double dt, int conserv_interp, int n,
                              double *x, double *z, int *phase, double *vx, double *vz) {
    if (!c || !u || !v || !x || !z || !phase || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (n == 0) return MDG_OK;
    advect_scratch &s = scratch_of(c);
    const size_t nu = (size_t)c->Nx * (c->Nz + 1), nv = (size_t)(c->Nx + 1) * c->Nz;
    const double *du, *dv;
    double *dx_, *dz_, *dvx, *dvz;
    int *dph;
    int rc;
    if ((rc = mdg_stage_in(c, u, nu, s.u, &du)) || (rc = mdg_stage_in(c, v, nv, s.v, &dv)) ||
        (rc = dev_view(c, x, (size_t)n, s.x, true, &dx_)) || (rc = dev_view(c, z, (size_t)n, s.z, true, &dz_)) ||
        (rc = dev_view(c, phase, (size_t)n, s.ph, true, &dph)) || (rc = dev_view(c, vx, (size_t)n, s.vx, true, &dvx)) ||
        (rc = dev_view(c, vz, (size_t)n, s.vz, true, &dvz)))
        return rc;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_advect_rk4, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, make_grid(c), du, dv, dt,
                       conserv_interp, n, dx_, dz_, dph, dvx, dvz);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if ((rc = copy_back(c, x, (const double *)dx_, (size_t)n)) || (rc = copy_back(c, z, (const double *)dz_, (size_t)n)) ||
        (rc = copy_back(c, phase, (const int *)dph, (size_t)n)) || (rc = copy_back(c, vx, (const double *)dvx, (size_t)n)) ||
        (rc = copy_back(c, vz, (const double *)dvz, (size_t)n)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return MDG_OK;
}
