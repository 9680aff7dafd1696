// Grid -> marker updates: Interp_Grid2P_centroids2 (ParticleRoutines.c:
// 1332-1580, non-periodic) and UpdateParticlePressure (RheologyParticles.c:
// 1621-1704) without subgrid diffusion. One thread per marker gathers the
// four centres around it (the centre field stays in L2); the marker streams
// x, z, phase in and the field (or P read-modify-write) out. Same arithmetic
// and order as the reference, built without contraction: bit-exact.
#include "mdg_internal.h"

namespace {

struct cgrid {
    int Nx, Nz;               // centres per row / column
    double dx, dz;
    const double *X, *Z;      // xvz, zvx (centre coordinates with one ghost each side)
    const signed char *tag;   // BCp
};

__device__ double g2p_cent_one(const cgrid &g, const double *__restrict__ F, double x, double z) {
    const int Nx = g.Nx, Nz = g.Nz;
    const double dx = g.dx, dz = g.dz;
    const int j = (int)ceil((x - g.X[0]) / dx) - 1;
    const int i = (int)ceil((z - g.Z[0]) / dz) - 1;
    int iSW = 0, iSE = 0, iNW = 0, iNE = 0;
    if (j > 0 && j < Nx && i > 0 && i < Nz) {
        iSW = (j - 1) + (i - 1) * Nx; iSE = iSW + 1; iNW = (j - 1) + i * Nx; iNE = iNW + 1;
    } else if (j == 0 && i > 0 && i < Nz) {          // west edge: copy along it
        iSW = iSE = (i - 1) * Nx; iNW = iNE = i * Nx;
    } else if (j == Nx && i > 0 && i < Nz) {         // east edge
        iSW = iSE = (j - 1) + (i - 1) * Nx; iNW = iNE = (j - 1) + i * Nx;
    } else if (i == 0 && j > 0 && j < Nx) {          // south edge
        iSW = iNW = j - 1; iSE = iNE = j;
    } else if (i == Nz && j > 0 && j < Nx) {         // north edge
        iSW = iNW = (j - 1) + (i - 1) * Nx; iSE = iNE = iSW + 1;
    } else if (i == 0 && j == 0) {
        iSW = iSE = iNW = iNE = 0;
    } else if (i == 0 && j == Nx) {
        iSW = iSE = iNW = iNE = j - 1;
    } else if (i == Nz && j == 0) {
        iSW = iSE = iNW = iNE = (i - 1) * Nx;
    } else if (i == Nz && j == Nx) {
        iSW = iSE = iNW = iNE = (j - 1) + (i - 1) * Nx;
    }
    // markers of the box always land in one of the cases above; clamp the
    // coordinate index for any other input so every read stays in range
    const int jj = j < 0 ? 0 : (j > Nx ? Nx : j), ii = i < 0 ? 0 : (i > Nz ? Nz : i);
    const double dxm = x - g.X[jj], dzm = z - g.Z[ii];
    const signed char *t = g.tag;
    const bool aSW = t[iSW] == 30 || t[iSW] == 31, aSE = t[iSE] == 30 || t[iSE] == 31;
    const bool aNW = t[iNW] == 30 || t[iNW] == 31, aNE = t[iNE] == 30 || t[iNE] == 31;
    double val = 0.0, sumW = 0.0;
    val += (1.0 - dxm / dx) * (1.0 - dzm / dz) * F[aSW ? iNW : iSW];
    sumW += (1.0 - dxm / dx) * (1.0 - dzm / dz);
    val += (dxm / dx) * (1.0 - dzm / dz) * F[aSE ? iNE : iSE];
    sumW += (dxm / dx) * (1.0 - dzm / dz);
    val += (1.0 - dxm / dx) * (dzm / dz) * F[aNW ? iSW : iNW];
    sumW += (1.0 - dxm / dx) * (dzm / dz);
    val += (dxm / dx) * (dzm / dz) * F[aNE ? iSE : iNE];
    sumW += (dxm / dx) * (dzm / dz);
    if (sumW > 1e-13) val /= sumW;
    return val;
}

__global__ void k_g2p_centroids(cgrid g, const double *__restrict__ F, int n, const double *__restrict__ x,
                                const double *__restrict__ z, const int *__restrict__ phase,
                                double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = (phase[k] != -1) ? g2p_cent_one(g, F, x[k], z[k]) : 0.0;
}

// Interp_Grid2P (ParticleRoutines.c:1223-1325) on the vertices
struct vgrid {
    int Nx, Nz;
    double dx, dz;
    const double *X, *Z;      // xg, zg
    const signed char *tag;   // BCg
};
__device__ double g2p_vert_one(const vgrid &g, const double *__restrict__ F, double xk, double zk) {
    const double dx = g.dx, dz = g.dz;
    int j = (int)ceil((xk - g.X[0]) / dx) - 1;
    if (j < 0) j = 0;
    if (j > g.Nx - 2) j = g.Nx - 2;
    int i = (int)ceil((zk - g.Z[0]) / dz) - 1;
    if (i < 0) i = 0;
    if (i > g.Nz - 2) i = g.Nz - 2;
    const double dxm = fabs(xk - g.X[j]), dzm = fabs(zk - g.Z[i]);
    const int iSW = j + i * g.Nx, iSE = iSW + 1, iNW = j + (i + 1) * g.Nx, iNE = iNW + 1;
    const signed char *t = g.tag;
    const bool aSW = t[iSW] == 30 || t[iSW] == 31, aSE = t[iSE] == 30 || t[iSE] == 31;
    const bool aNW = t[iNW] == 30 || t[iNW] == 31, aNE = t[iNE] == 30 || t[iNE] == 31;
    double val = 0.0;
    val += (1.0 - dxm / dx) * (1.0 - dzm / dz) * F[aSW ? iNW : iSW];
    val += (dxm / dx) * (1.0 - dzm / dz) * F[aSE ? iNE : iSE];
    val += (1.0 - dxm / dx) * (dzm / dz) * F[aNW ? iSW : iNW];
    val += (dxm / dx) * (dzm / dz) * F[aNE ? iSE : iNE];
    return val;
}

__global__ void k_g2p_vertices(vgrid g, const double *__restrict__ F, int n, const double *__restrict__ x,
                               const double *__restrict__ z, const int *__restrict__ phase,
                               double *__restrict__ out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = (phase[k] != -1) ? g2p_vert_one(g, F, x[k], z[k]) : 0.0;
}

// grid stress increments of UpdateParticleStress (RheologyParticles.c:1929-1939)
__global__ void k_stress_increments(int nc, int nv, const signed char *__restrict__ tp,
                                    const signed char *__restrict__ tg, const double *__restrict__ sxx,
                                    const double *__restrict__ szz, const double *__restrict__ sxz,
                                    const double *__restrict__ sxx0, const double *__restrict__ szz0,
                                    const double *__restrict__ sxz0, double *__restrict__ dxx,
                                    double *__restrict__ dzz, double *__restrict__ dxz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) {
        const bool on = tp[i] != 30 && tp[i] != 31;
        dxx[i] = on ? sxx[i] - sxx0[i] : 0.0;
        dzz[i] = on ? szz[i] - szz0[i] : 0.0;
    } else if (i < nc + nv) {
        const int j = i - nc;
        dxz[j] = (tg[j] != 30) ? sxz[j] - sxz0[j] : 0.0;
    }
}
__global__ void k_add_stress(cgrid gc, vgrid gv, const double *__restrict__ dxx, const double *__restrict__ dzz,
                             const double *__restrict__ dxz, int n, const double *__restrict__ x,
                             const double *__restrict__ z, const int *__restrict__ phase, double *__restrict__ mxx,
                             double *__restrict__ mzz, double *__restrict__ mxz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const bool on = phase[k] != -1;
    const double xk = x[k], zk = z[k];
    mxx[k] = mxx[k] + (on ? g2p_cent_one(gc, dxx, xk, zk) : 0.0);
    mzz[k] = mzz[k] + (on ? g2p_cent_one(gc, dzz, xk, zk) : 0.0);
    mxz[k] = mxz[k] + (on ? g2p_vert_one(gv, dxz, xk, zk) : 0.0);
}

__global__ void k_pressure_increment(int nc, const signed char *__restrict__ tag, const double *__restrict__ p_in,
                                     const double *__restrict__ p0, double *__restrict__ dp) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) dp[c] = (tag[c] != 30) ? (p_in[c] - p0[c]) : 0.0;
}

__global__ void k_add_g2p(cgrid g, const double *__restrict__ F, int n, const double *__restrict__ x,
                          const double *__restrict__ z, const int *__restrict__ phase, double *__restrict__ P) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double inc = (phase[k] != -1) ? g2p_cent_one(g, F, x[k], z[k]) : 0.0;
    P[k] = P[k] + inc;
}

// AccumulatedStrainII (RheologyParticles.c:346-649): per-centre increments
// dt * eII in the reference's order (the total as the sum of the seven
// mechanisms pl, pwl, exp, lin, gbs, el, cst), into 8 arrays of nc; a
// negative total sets *bad to the smallest such cell + 1 (the reference
// exits there).
struct strain_rates7 {
    const double *e[7];      // eII_pl, eII_pwl, eII_exp, eII_lin, eII_gbs, eII_el, eII_cst (NULL = 0)
    const double *div_pl;    // div_u_pl (NULL = 0)
};
__global__ void k_strain_increments(int nc, double dt, strain_rates7 r, double *__restrict__ inc, int *__restrict__ bad) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double e[7];
    for (int q = 0; q < 7; q++) e[q] = r.e[q] ? r.e[q][c] : 0.0;
    const double tot = dt * (e[0] + e[1] + e[2] + e[3] + e[4] + e[5] + e[6]);
    if (tot < 0) atomicMin(bad, c + 1);
    inc[c] = tot;
    inc[1 * (size_t)nc + c] = dt * e[5];
    inc[2 * (size_t)nc + c] = dt * e[0];
    inc[3 * (size_t)nc + c] = dt * (r.div_pl ? r.div_pl[c] : 0.0);
    inc[4 * (size_t)nc + c] = dt * e[1];
    inc[5 * (size_t)nc + c] = dt * e[2];
    inc[6 * (size_t)nc + c] = dt * e[3];
    inc[7 * (size_t)nc + c] = dt * e[4];
}

// Marker part: the four centres around the marker (the cell index clamped to
// the centre grid, BCp 30/31 centres left out, weights renormalised), each
// increment interpolated and added to the marker's strain (NULL arrays are
// skipped; inactive markers add zero, as the reference's loop does).
struct strain_out8 {
    double *s[8];            // strain, el, pl, pl_vol, pwl, exp, lin, gbs
};
__global__ void k_accumulate_strain(int Ncx, int Ncz, double dx, double dz, const double *__restrict__ xc,
                                    const double *__restrict__ zc, const signed char *__restrict__ tag,
                                    const double *__restrict__ inc, int n, const double *__restrict__ x,
                                    const double *__restrict__ z, const int *__restrict__ phase, strain_out8 out) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const size_t nc = (size_t)Ncx * Ncz;
    double dE[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (phase[k] != -1) {
        const double xk = x[k], zk = z[k];
        int j = (int)ceil((xk - xc[0]) / dx) - 1;
        j = j < 0 ? 0 : (j > Ncx - 2 ? Ncx - 2 : j);
        int i = (int)ceil((zk - zc[0]) / dz) - 1;
        i = i < 0 ? 0 : (i > Ncz - 2 ? Ncz - 2 : i);
        const double dxm = xk - xc[j], dzm = zk - zc[i];
        const int nd[4] = {j + i * Ncx, j + i * Ncx + 1, j + (i + 1) * Ncx, j + (i + 1) * Ncx + 1};
        const double w[4] = {(1.0 - dxm / dx) * (1.0 - dzm / dz), (dxm / dx) * (1.0 - dzm / dz),
                             (1.0 - dxm / dx) * (dzm / dz), (dxm / dx) * (dzm / dz)};
        double sumW = 0.0;
        for (int q = 0; q < 4; q++) {
            const signed char t = tag[nd[q]];
            if (t == 30 || t == 31) continue;
            for (int f = 0; f < 8; f++) dE[f] += w[q] * inc[f * nc + nd[q]];
            sumW += w[q];
        }
        if (sumW > 1e-13)
            for (int f = 0; f < 8; f++) dE[f] /= sumW;
    }
    for (int f = 0; f < 8; f++)
        if (out.s[f]) out.s[f][k] = out.s[f][k] + dE[f];
}

// Stress rotation of UpdateParticleStress (RheologyParticles.c:1965-1988)
__global__ void k_rotate_stress(int mode, double dt, int n, const int *__restrict__ phase, const double *__restrict__ w,
                                double *__restrict__ sxx, double *__restrict__ szz, double *__restrict__ sxz) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n || phase[k] == -1) return;
    const double txx = sxx[k], tzz = szz[k], txz = sxz[k], wk = w[k];
    if (mode == 1) {   // Jaumann rate
        sxx[k] = txx - dt * 2.0 * txz * wk;
        szz[k] = tzz - -dt * 2.0 * txz * wk;
        sxz[k] = txz - dt * (tzz - txx) * wk;
    } else {           // rotation by dt * w
        const double a = dt * wk, c = cos(a), s = sin(a);
        sxx[k] = (txx * c + txz * s) * c + (txz * c + tzz * s) * s;
        szz[k] = -(-txx * s + txz * c) * s + (-txz * s + tzz * c) * c;
        sxz[k] = -(txx * c + txz * s) * s + (txz * c + tzz * s) * c;
    }
}

cgrid make_cgrid(const mdg_ctx *c) {
    return cgrid{c->Nx - 1, c->Nz - 1, c->dx, c->dz, c->xvz.p, c->zvx.p, c->BCp_t.p};
}

// Subgrid relaxation of UpdateParticleStress / UpdateParticlePressure
// (RheologyParticles.c:1845-1850, :1660-1666): on active markers
//   ms = -(m - m0) (1 - exp(-d dt / (eta_m / coef[phase]))),  d = 1.
struct relax3 {
    const double *m[3], *m0[3];
    double *ms[3];
};
__global__ void k_subgrid_relax(int n, int nf, double dt, const double *__restrict__ coef, const int *__restrict__ phase,
                                const double *__restrict__ etam, relax3 a) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int p = phase[k];
    if (p == -1) {
        for (int q = 0; q < nf; q++) a.ms[q][k] = 0.0;
        return;
    }
    const double d = 1.0, trel = etam[k] / coef[p];
    for (int q = 0; q < nf; q++) a.ms[q][k] = -(a.m[q][k] - a.m0[q][k]) * (1.0 - exp(-d * dt / trel));
}
// remaining grid increments (tau - tau0) - dtau_s on non-air nodes
__global__ void k_stress_remaining(int nc, int nv, const signed char *__restrict__ tp, const signed char *__restrict__ tg,
                                   const double *__restrict__ sxx, const double *__restrict__ szz,
                                   const double *__restrict__ sxz, const double *__restrict__ sxx0,
                                   const double *__restrict__ szz0, const double *__restrict__ sxz0,
                                   const double *__restrict__ gxx, const double *__restrict__ gzz,
                                   const double *__restrict__ gxz, double *__restrict__ rxx, double *__restrict__ rzz,
                                   double *__restrict__ rxz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nc) {
        const bool on = tp[i] != 30 && tp[i] != 31;
        rxx[i] = on ? (sxx[i] - sxx0[i]) - gxx[i] : 0.0;
        rzz[i] = on ? (szz[i] - szz0[i]) - gzz[i] : 0.0;
    } else if (i < nc + nv) {
        const int j = i - nc;
        rxz[j] = (tg[j] != 30) ? (sxz[j] - sxz0[j]) - gxz[j] : 0.0;
    }
}
// old pressure on non-air cells, and later the remaining increment dp - dP_s
__global__ void k_masked(int nc, const signed char *__restrict__ tag, const double *__restrict__ a,
                         const double *__restrict__ b, double *__restrict__ out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc) out[c] = (tag[c] != 30) ? (b ? a[c] - b[c] : a[c]) : 0.0;
}
// m = m + ms + mr on active markers
__global__ void k_add_two(int n, const int *__restrict__ phase, const double *__restrict__ ms,
                          const double *__restrict__ mr, double *__restrict__ m) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n && phase[k] != -1) m[k] = m[k] + ms[k] + mr[k];
}

}  // namespace

// staging buffers for host callers, one set per context (freed with it)
struct g2p_scratch {
    dbuf<double> F, x, z, out, p0, dp;
    dbuf<int> ph;
    dbuf<double> in[6], inc, m[3];   // stress update: grid inputs, increments, marker stresses
    dbuf<double> e[8], ms[8];        // accumulated strain: grid rates, marker strains (host callers)
    dbuf<double> sub[10], gsub[6], tab;   // subgrid diffusion: marker and grid temporaries, phase table
};

static g2p_scratch &g2p_of(mdg_ctx *c) {
    if (!c->g2p) c->g2p = new g2p_scratch();
    return *c->g2p;
}

void mdg_g2p_release(mdg_ctx *c) {
    if (!c || !c->g2p) return;
    g2p_scratch *s = c->g2p;
    s->F.release(); s->x.release(); s->z.release(); s->out.release(); s->p0.release(); s->dp.release();
    s->ph.release(); s->inc.release();
    for (auto &b : s->in) b.release();
    for (auto &b : s->m) b.release();
    for (auto &b : s->e) b.release();
    for (auto &b : s->ms) b.release();
    for (auto &b : s->sub) b.release();
    for (auto &b : s->gsub) b.release();
    s->tab.release();
    delete s;
    c->g2p = nullptr;
}

static int timed_end(mdg_ctx *c) {
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return MDG_OK;
}

extern "C" int mdg_g2p_centroids(mdg_ctx *c, const double *F, int n, const double *x, const double *z,
                                 const int *phase, double *out) {
    if (!c || !F || !x || !z || !phase || !out || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (n == 0) return MDG_OK;
    g2p_scratch &s = g2p_of(c);
    const size_t nc = (size_t)(c->Nx - 1) * (c->Nz - 1);
    const double *dF, *dx_, *dz_;
    const int *dph;
    int rc;
    if ((rc = mdg_stage_in(c, F, nc, s.F, &dF)) || (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) ||
        (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) || (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    double *dout = out;
    if (!mdg_is_device_ptr(out)) {
        HIPCHK(c, s.out.reserve((size_t)n));
        dout = s.out.p;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_g2p_centroids, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, make_cgrid(c), dF, n, dx_,
                       dz_, dph, dout);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (dout != out) HIPCHK(c, hipMemcpyAsync(out, dout, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

extern "C" int mdg_g2p_vertices(mdg_ctx *c, const double *F, int n, const double *x, const double *z,
                                const int *phase, double *out) {
    if (!c || !F || !x || !z || !phase || !out || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (n == 0) return MDG_OK;
    g2p_scratch &s = g2p_of(c);
    const double *dF, *dx_, *dz_;
    const int *dph;
    int rc;
    if ((rc = mdg_stage_in(c, F, (size_t)c->Nx * c->Nz, s.F, &dF)) || (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) ||
        (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) || (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    double *dout = out;
    if (!mdg_is_device_ptr(out)) {
        HIPCHK(c, s.out.reserve((size_t)n));
        dout = s.out.p;
    }
    const vgrid g = {c->Nx, c->Nz, c->dx, c->dz, c->xg.p, c->zg.p, c->BCg_t.p};
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_g2p_vertices, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, g, dF, n, dx_, dz_, dph, dout);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (dout != out) HIPCHK(c, hipMemcpyAsync(out, dout, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

extern "C" int mdg_update_particle_pressure(mdg_ctx *c, const double *p_in, const double *p0_n, double *dp, int n,
                                            const double *x, const double *z, const int *phase, double *P) {
    if (!c || !p_in || !p0_n || !x || !z || !phase || !P || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    g2p_scratch &s = g2p_of(c);
    const int nc = (c->Nx - 1) * (c->Nz - 1);
    const double *dpin, *dp0, *dx_, *dz_;
    const int *dph;
    int rc;
    if ((rc = mdg_stage_in(c, p_in, (size_t)nc, s.F, &dpin)) || (rc = mdg_stage_in(c, p0_n, (size_t)nc, s.p0, &dp0)) ||
        (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) || (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) ||
        (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    double *ddp = dp;
    if (!dp || !mdg_is_device_ptr(dp)) {
        HIPCHK(c, s.dp.reserve((size_t)nc));
        ddp = s.dp.p;
    }
    double *dP = P;
    if (!mdg_is_device_ptr(P)) {
        HIPCHK(c, s.out.reserve((size_t)(n > 0 ? n : 1)));
        if (n > 0) HIPCHK(c, hipMemcpyAsync(s.out.p, P, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        dP = s.out.p;
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_pressure_increment, dim3(mdg_grid1d(nc, 256)), dim3(256), 0, c->stream, nc,
                       (const signed char *)c->BCp_t.p, dpin, dp0, ddp);
    if (n > 0)
        hipLaunchKernelGGL(k_add_g2p, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, make_cgrid(c), (const double *)ddp,
                           n, dx_, dz_, dph, dP);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (dP != P && n > 0) HIPCHK(c, hipMemcpyAsync(P, dP, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    if (dp && ddp != dp) HIPCHK(c, hipMemcpyAsync(dp, ddp, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

extern "C" int mdg_update_particle_stress(mdg_ctx *c, const double *sxxd, const double *szzd, const double *sxz,
                                          const double *sxxd0, const double *szzd0, const double *sxz0, int n,
                                          const double *x, const double *z, const int *phase, double *msxxd,
                                          double *mszzd, double *msxz) {
    if (!c || !sxxd || !szzd || !sxz || !sxxd0 || !szzd0 || !sxz0 || !x || !z || !phase || !msxxd || !mszzd ||
        !msxz || n < 0)
        return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (n == 0) return MDG_OK;
    g2p_scratch &s = g2p_of(c);
    const int nc = (c->Nx - 1) * (c->Nz - 1), nv = c->Nx * c->Nz;
    const double *src[6] = {sxxd, szzd, sxz, sxxd0, szzd0, sxz0}, *d[6];
    const size_t len[6] = {(size_t)nc, (size_t)nc, (size_t)nv, (size_t)nc, (size_t)nc, (size_t)nv};
    int rc;
    for (int q = 0; q < 6; q++)
        if ((rc = mdg_stage_in(c, src[q], len[q], s.in[q], &d[q]))) return rc;
    const double *dx_, *dz_;
    const int *dph;
    if ((rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) || (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) ||
        (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    HIPCHK(c, s.inc.reserve((size_t)(2 * nc + nv)));
    double *dxx = s.inc.p, *dzz = dxx + nc, *dxz = dzz + nc;
    double *user[3] = {msxxd, mszzd, msxz}, *dm[3];
    for (int q = 0; q < 3; q++) {
        if (mdg_is_device_ptr(user[q])) {
            dm[q] = user[q];
        } else {
            HIPCHK(c, s.m[q].reserve((size_t)n));
            HIPCHK(c, hipMemcpyAsync(s.m[q].p, user[q], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            dm[q] = s.m[q].p;
        }
    }
    const vgrid gv = {c->Nx, c->Nz, c->dx, c->dz, c->xg.p, c->zg.p, c->BCg_t.p};
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_stress_increments, dim3(mdg_grid1d(nc + nv, 256)), dim3(256), 0, c->stream, nc, nv,
                       (const signed char *)c->BCp_t.p, (const signed char *)c->BCg_t.p, d[0], d[1], d[2], d[3], d[4],
                       d[5], dxx, dzz, dxz);
    hipLaunchKernelGGL(k_add_stress, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, make_cgrid(c), gv,
                       (const double *)dxx, (const double *)dzz, (const double *)dxz, n, dx_, dz_, dph, dm[0], dm[1],
                       dm[2]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    for (int q = 0; q < 3; q++)
        if (dm[q] != user[q]) HIPCHK(c, hipMemcpyAsync(user[q], dm[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

extern "C" int mdg_accumulated_strain(mdg_ctx *c, double dt, const double *const *eII, const double *div_u_pl, int n,
                                      const double *x, const double *z, const int *phase, double *const *strain) {
    if (!c || !eII || !strain || !x || !z || !phase || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    g2p_scratch &s = g2p_of(c);
    const int Ncx = c->Nx - 1, Ncz = c->Nz - 1, nc = Ncx * Ncz;
    strain_rates7 r{};
    int rc;
    for (int q = 0; q < 7; q++)
        if ((rc = mdg_stage_in(c, eII[q], (size_t)nc, s.e[q], &r.e[q]))) return rc;
    if ((rc = mdg_stage_in(c, div_u_pl, (size_t)nc, s.e[7], &r.div_pl))) return rc;
    const double *dx_, *dz_;
    const int *dph;
    if ((rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) || (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) ||
        (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    strain_out8 out{};
    for (int f = 0; f < 8; f++) {
        if (!strain[f] || mdg_is_device_ptr(strain[f])) {
            out.s[f] = strain[f];
        } else {
            HIPCHK(c, s.ms[f].reserve((size_t)(n > 0 ? n : 1)));
            if (n > 0) HIPCHK(c, hipMemcpyAsync(s.ms[f].p, strain[f], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            out.s[f] = s.ms[f].p;
        }
    }
    HIPCHK(c, s.inc.reserve(8 * (size_t)nc));
    int *bad = c->flag_dev.p, h_bad = 0;
    HIPCHK(c, hipMemsetAsync(bad, 0x7f, sizeof(int), c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_strain_increments, dim3(mdg_grid1d(nc, 256)), dim3(256), 0, c->stream, nc, dt, r, s.inc.p, bad);
    HIPCHK(c, hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h_bad != 0x7f7f7f7f)
        return mdg_fail(c, MDG_ERR_ARG, "negative strain increment at cell %d (AccumulatedStrainII exits there)", h_bad - 1);
    if (n > 0)
        hipLaunchKernelGGL(k_accumulate_strain, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, Ncx, Ncz, c->dx, c->dz,
                           (const double *)c->xc.p, (const double *)c->zc.p, (const signed char *)c->BCp_t.p,
                           (const double *)s.inc.p, n, dx_, dz_, dph, out);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    for (int f = 0; f < 8; f++)
        if (strain[f] && out.s[f] != strain[f] && n > 0)
            HIPCHK(c, hipMemcpyAsync(strain[f], out.s[f], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

extern "C" int mdg_rotate_particle_stress(mdg_ctx *c, int mode, double dt, int n, const int *phase, const double *wxzm,
                                          double *sxxd, double *szzd, double *sxz) {
    if (!c || !phase || !wxzm || !sxxd || !szzd || !sxz || n < 0) return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (mode < 0 || mode > 2) return mdg_fail(c, MDG_ERR_ARG, "stress_rotation must be 0, 1 or 2");
    if (mode == 0 || n == 0) return MDG_OK;
    g2p_scratch &s = g2p_of(c);
    const int *dph;
    const double *dw;
    int rc;
    if ((rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)) || (rc = mdg_stage_in(c, wxzm, (size_t)n, s.x, &dw)))
        return rc;
    double *user[3] = {sxxd, szzd, sxz}, *dm[3];
    for (int q = 0; q < 3; q++) {
        if (mdg_is_device_ptr(user[q])) {
            dm[q] = user[q];
        } else {
            HIPCHK(c, s.m[q].reserve((size_t)n));
            HIPCHK(c, hipMemcpyAsync(s.m[q].p, user[q], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            dm[q] = s.m[q].p;
        }
    }
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_rotate_stress, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, mode, dt, n, dph, dw, dm[0],
                       dm[1], dm[2]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    for (int q = 0; q < 3; q++)
        if (dm[q] != user[q]) HIPCHK(c, hipMemcpyAsync(user[q], dm[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    return timed_end(c);
}

// Marker arrays: device pointers used in place, host arrays staged in and
// copied back by the caller of this helper.
static int stage_markers_rw(mdg_ctx *c, dbuf<double> *tmp, double *const *user, int nq, int n, double **dev) {
    for (int q = 0; q < nq; q++) {
        if (mdg_is_device_ptr(user[q])) {
            dev[q] = user[q];
        } else {
            HIPCHK(c, tmp[q].reserve((size_t)(n > 0 ? n : 1)));
            if (n > 0) HIPCHK(c, hipMemcpyAsync(tmp[q].p, user[q], (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            dev[q] = tmp[q].p;
        }
    }
    return MDG_OK;
}

struct local_timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~local_timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

extern "C" int mdg_update_particle_stress_subgrid(mdg_ctx *c, double dt, const double *G_phase, int nph,
                                                  int interp_stencil, const double *sxxd, const double *szzd,
                                                  const double *sxz, const double *sxxd0, const double *szzd0,
                                                  const double *sxz0, const double *eta_phys_s, int n,
                                                  const double *x, const double *z, const int *phase, double *msxxd,
                                                  double *mszzd, double *msxz) {
    if (!c || !G_phase || nph < 1 || !sxxd || !szzd || !sxz || !sxxd0 || !szzd0 || !sxz0 || !eta_phys_s || !x || !z ||
        !phase || !msxxd || !mszzd || !msxz || n < 0)
        return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (c->n_markers != n)
        return mdg_fail(c, MDG_ERR_STATE, "the subgrid step needs these markers set (mdg_markers_set): %d set, %d given",
                        c->n_markers, n);
    if (n == 0) return MDG_OK;
    g2p_scratch &s = g2p_of(c);
    const int nc = (c->Nx - 1) * (c->Nz - 1), nv = c->Nx * c->Nz;
    const double *src[7] = {sxxd, szzd, sxz, sxxd0, szzd0, sxz0, eta_phys_s}, *d[7];
    const size_t len[7] = {(size_t)nc, (size_t)nc, (size_t)nv, (size_t)nc, (size_t)nc, (size_t)nv, (size_t)nv};
    int rc;
    for (int q = 0; q < 6; q++)
        if ((rc = mdg_stage_in(c, src[q], len[q], s.in[q], &d[q]))) return rc;
    if ((rc = mdg_stage_in(c, src[6], len[6], s.gsub[5], &d[6]))) return rc;
    const double *dG, *dx_, *dz_;
    const int *dph;
    if ((rc = mdg_stage_in(c, G_phase, (size_t)nph, s.tab, &dG)) || (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) ||
        (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) || (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    double *user[3] = {msxxd, mszzd, msxz}, *dm[3];
    if ((rc = stage_markers_rw(c, s.m, user, 3, n, dm))) return rc;
    for (int q = 0; q < 10; q++) HIPCHK(c, s.sub[q].reserve((size_t)n));
    double *m0[3] = {s.sub[0].p, s.sub[1].p, s.sub[2].p}, *etam = s.sub[3].p;
    double *ms[3] = {s.sub[4].p, s.sub[5].p, s.sub[6].p}, *mr[3] = {s.sub[7].p, s.sub[8].p, s.sub[9].p};
    const size_t gl[5] = {(size_t)nc, (size_t)nc, (size_t)nv, (size_t)nc, (size_t)nc};
    for (int q = 0; q < 5; q++) HIPCHK(c, s.gsub[q].reserve(gl[q]));
    HIPCHK(c, s.inc.reserve((size_t)nv));
    double *gs[3] = {s.gsub[0].p, s.gsub[1].p, s.gsub[2].p}, *gr[3] = {s.gsub[3].p, s.gsub[4].p, s.inc.p};
    const cgrid gc = make_cgrid(c);
    const vgrid gv = {c->Nx, c->Nz, c->dx, c->dz, c->xg.p, c->zg.p, c->BCg_t.p};
    local_timer t;
    HIPCHK(c, hipEventCreate(&t.a));
    HIPCHK(c, hipEventCreate(&t.b));
    HIPCHK(c, hipEventRecord(t.a, c->stream));
    const dim3 gm(mdg_grid1d(n, 256)), bk(256);
    // old stresses and the vertex viscosity at the markers
    hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, d[3], n, dx_, dz_, dph, m0[0]);
    hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, d[4], n, dx_, dz_, dph, m0[1]);
    hipLaunchKernelGGL(k_g2p_vertices, gm, bk, 0, c->stream, gv, d[5], n, dx_, dz_, dph, m0[2]);
    hipLaunchKernelGGL(k_g2p_vertices, gm, bk, 0, c->stream, gv, d[6], n, dx_, dz_, dph, etam);
    relax3 ra{{dm[0], dm[1], dm[2]}, {m0[0], m0[1], m0[2]}, {ms[0], ms[1], ms[2]}};
    hipLaunchKernelGGL(k_subgrid_relax, gm, bk, 0, c->stream, n, 3, dt, dG, dph, (const double *)etam, ra);
    HIPCHK(c, hipGetLastError());
    // subgrid increments onto the grid (P2Mastah, flag 1, arithmetic)
    for (int q = 0; q < 3; q++) {
        HIPCHK(c, hipMemsetAsync(gs[q], 0, (q < 2 ? (size_t)nc : (size_t)nv) * 8, c->stream));
        if ((rc = mdg_p2mastah(c, ms[q], n, gs[q], q < 2 ? c->BCp_t.p : c->BCg_t.p, 1, 0, 0, q < 2 ? 1 : 0,
                               interp_stencil, 0, nullptr)))
            return rc;
    }
    hipLaunchKernelGGL(k_stress_remaining, dim3(mdg_grid1d(nc + nv, 256)), bk, 0, c->stream, nc, nv,
                       (const signed char *)c->BCp_t.p, (const signed char *)c->BCg_t.p, d[0], d[1], d[2], d[3], d[4],
                       d[5], (const double *)gs[0], (const double *)gs[1], (const double *)gs[2], gr[0], gr[1], gr[2]);
    hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, (const double *)gr[0], n, dx_, dz_, dph, mr[0]);
    hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, (const double *)gr[1], n, dx_, dz_, dph, mr[1]);
    hipLaunchKernelGGL(k_g2p_vertices, gm, bk, 0, c->stream, gv, (const double *)gr[2], n, dx_, dz_, dph, mr[2]);
    for (int q = 0; q < 3; q++)
        hipLaunchKernelGGL(k_add_two, gm, bk, 0, c->stream, n, dph, (const double *)ms[q], (const double *)mr[q], dm[q]);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(t.b, c->stream));
    for (int q = 0; q < 3; q++)
        if (dm[q] != user[q]) HIPCHK(c, hipMemcpyAsync(user[q], dm[q], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms_ = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms_, t.a, t.b));
    c->last_ms = ms_;
    return MDG_OK;
}

extern "C" int mdg_update_particle_pressure_subgrid(mdg_ctx *c, double dt, const double *beta_phase, int nph,
                                                    const double *p_in, const double *p0_n, const double *eta_phys_n,
                                                    double *dp, int n, const double *x, const double *z,
                                                    const int *phase, double *P) {
    if (!c || !beta_phase || nph < 1 || !p_in || !p0_n || !eta_phys_n || !x || !z || !phase || !P || n < 0)
        return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->have_bc) return mdg_fail(c, MDG_ERR_STATE, "mdg_set_bc has not been called");
    if (c->n_markers != n)
        return mdg_fail(c, MDG_ERR_STATE, "the subgrid step needs these markers set (mdg_markers_set): %d set, %d given",
                        c->n_markers, n);
    g2p_scratch &s = g2p_of(c);
    const int nc = (c->Nx - 1) * (c->Nz - 1);
    const double *dpin, *dp0, *deta, *dB, *dx_, *dz_;
    const int *dph;
    int rc;
    if ((rc = mdg_stage_in(c, p_in, (size_t)nc, s.F, &dpin)) || (rc = mdg_stage_in(c, p0_n, (size_t)nc, s.p0, &dp0)) ||
        (rc = mdg_stage_in(c, eta_phys_n, (size_t)nc, s.gsub[5], &deta)) ||
        (rc = mdg_stage_in(c, beta_phase, (size_t)nph, s.tab, &dB)) || (rc = mdg_stage_in(c, x, (size_t)n, s.x, &dx_)) ||
        (rc = mdg_stage_in(c, z, (size_t)n, s.z, &dz_)) || (rc = mdg_stage_in(c, phase, (size_t)n, s.ph, &dph)))
        return rc;
    double *ddp = dp;
    if (!dp || !mdg_is_device_ptr(dp)) {
        HIPCHK(c, s.dp.reserve((size_t)nc));
        ddp = s.dp.p;
    }
    double *user[1] = {P}, *dP[1];
    if ((rc = stage_markers_rw(c, s.m, user, 1, n, dP))) return rc;
    for (int q = 0; q < 4; q++) HIPCHK(c, s.sub[q].reserve((size_t)(n > 0 ? n : 1)));
    double *etam = s.sub[0].p, *Pm0 = s.sub[1].p, *dPms = s.sub[2].p, *dPmr = s.sub[3].p;
    for (int q = 0; q < 3; q++) HIPCHK(c, s.gsub[q].reserve((size_t)nc));
    double *Pg0 = s.gsub[0].p, *dPgs = s.gsub[1].p, *dPgr = s.gsub[2].p;
    const cgrid gc = make_cgrid(c);
    local_timer t;
    HIPCHK(c, hipEventCreate(&t.a));
    HIPCHK(c, hipEventCreate(&t.b));
    HIPCHK(c, hipEventRecord(t.a, c->stream));
    const dim3 gm(mdg_grid1d(n > 0 ? n : 1, 256)), gcl(mdg_grid1d(nc, 256)), bk(256);
    const signed char *tag = c->BCp_t.p;
    hipLaunchKernelGGL(k_pressure_increment, gcl, bk, 0, c->stream, nc, tag, dpin, dp0, ddp);
    if (n > 0) {
        hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, deta, n, dx_, dz_, dph, etam);
        hipLaunchKernelGGL(k_masked, gcl, bk, 0, c->stream, nc, tag, dp0, (const double *)nullptr, Pg0);
        hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, (const double *)Pg0, n, dx_, dz_, dph, Pm0);
        relax3 ra{{dP[0], nullptr, nullptr}, {Pm0, nullptr, nullptr}, {dPms, nullptr, nullptr}};
        hipLaunchKernelGGL(k_subgrid_relax, gm, bk, 0, c->stream, n, 1, dt, dB, dph, (const double *)etam, ra);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemsetAsync(dPgs, 0, (size_t)nc * 8, c->stream));
        if ((rc = mdg_p2mastah(c, dPms, n, dPgs, tag, 1, 0, 0, 1, 1, 0, nullptr))) return rc;
        hipLaunchKernelGGL(k_masked, gcl, bk, 0, c->stream, nc, tag, (const double *)ddp, (const double *)dPgs, dPgr);
        hipLaunchKernelGGL(k_g2p_centroids, gm, bk, 0, c->stream, gc, (const double *)dPgr, n, dx_, dz_, dph, dPmr);
        hipLaunchKernelGGL(k_add_two, gm, bk, 0, c->stream, n, dph, (const double *)dPms, (const double *)dPmr, dP[0]);
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(t.b, c->stream));
    if (dP[0] != P && n > 0) HIPCHK(c, hipMemcpyAsync(P, dP[0], (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    if (dp && ddp != dp) HIPCHK(c, hipMemcpyAsync(dp, ddp, (size_t)nc * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms_ = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms_, t.a, t.b));
    c->last_ms = ms_;
    return MDG_OK;
}
