// Implicit energy step on the GPU: the matrix of EnergyDirectSolve
// (ThermalRoutines.c:185-390, theta = 1, transient, no air cells) applied
// matrix-free from the cell-face conductivities, solved with the
// Jacobi-preconditioned CG of SolveThermalPCG (ThermalSolver.c:338-480)
// started from the current temperature.
//
// Per iteration four launches on the context's stream:
//   apply  : p = z + beta p_old on the cell and its four neighbours (the
//            previous iteration's direction update, folded in here), q = A p,
//            block partials of p.q;
//   fin1   : alpha = rz / p.q (breakdown -2);
//   update : x += alpha p, r -= alpha q, z = M^-1 r, block partials of r.r
//            and r.z;
//   fin2   : ||r|| / ||r0|| test, breakdown -3, beta, iteration limit.
// The control block lives in device memory; once it says "done" every later
// launch returns at once, so the host queues iterations in batches and
// reads the control block between batches only.  Products and sums within a
// cell follow the reference's CSR row order (south, west, centre, east,
// north), so only the dot-product reductions round differently.
#include "mdg_internal.h"

namespace {

constexpr int TB = 256;

struct th_ctl {
    double rz, r0, alpha, beta, rel_tol;
    int done, its, it, max_its;
};

struct th_geom {
    int ncx, ncz, nx, nxvz;
    double idx2, idz2, dx;
};

// Block b of nb takes a contiguous share of the cell range per XCD (blocks
// are dispatched round-robin over the 8 XCDs), so the rows above and below a
// block's cells are in the same L2 as the neighbouring blocks that read them.
__device__ inline int xcd_block(int b, int nb) {
    const int xcd = b & 7, per = nb >> 3, rem = nb & 7;
    return xcd * per + min(xcd, rem) + (b >> 3);
}

__device__ inline double block_sum(double v, double *s) {
    s[threadIdx.x] = v;
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
        __syncthreads();
    }
    return s[0];
}

// Diagonal and boundary right-hand side of cell (k, l), in the order of the
// reference's assembly.
__device__ inline void th_diag(const th_geom &g, double rhoCp_dt, const double *__restrict__ kx,
                               const double *__restrict__ kz, const signed char *__restrict__ BT,
                               const double *__restrict__ BV, int k, int l, double *d, double *bbc) {
    const int c0 = (k + 1) + (l + 1) * (g.ncx + 2), c1 = k + (l + 1) * g.nx, c3 = k + l * g.nxvz + 1;
    const double AW = kx[c1], AE = kx[c1 + 1], AS = kz[c3], AN = kz[c3 + g.nxvz];
    double val = rhoCp_dt, b = 0.0;
    if (l > 0) val += AS * g.idz2;
    if (k > 0) val += AW * g.idx2;
    if (k < g.ncx - 1) val += AE * g.idx2;
    if (l < g.ncz - 1) val += AN * g.idz2;
    if (l == 0 && BT[c0 - (g.ncx + 2)] == 1) {
        val += 2.0 * AS * g.idz2;
        b += 2.0 * AS * g.idz2 * BV[c0 - (g.ncx + 2)];
    }
    if (l == 0 && BT[c0 - (g.ncx + 2)] == 0) b += -1.0 / g.dx * BV[c0 - (g.ncx + 2)];
    if (k == 0 && BT[c0 - 1] == 1) {
        val += 2.0 * AW * g.idx2;
        b += 2.0 * AW * g.idx2 * BV[c0 - 1];
    }
    if (k == g.ncx - 1 && BT[c0 + 1] == 1) {
        val += 2.0 * AE * g.idx2;
        b += 2.0 * AE * g.idx2 * BV[c0 + 1];
    }
    if (l == g.ncz - 1 && BT[c0 + (g.ncx + 2)] == 1) {
        val += 2.0 * AN * g.idz2;
        b += 2.0 * AN * g.idz2 * BV[c0 + (g.ncx + 2)];
    }
    *d = val;
    *bbc = b;
}

// Row c2 = (k, l) of A times v, with P(j) giving v at cell j.
template <typename F>
__device__ inline double th_row(const th_geom &g, const double *__restrict__ kx, const double *__restrict__ kz,
                                double d, int k, int l, int c2, F P) {
    const int c1 = k + (l + 1) * g.nx, c3 = k + l * g.nxvz + 1;
    double s = 0.0;
    if (l > 0) s += -kz[c3] * g.idz2 * P(c2 - g.ncx);
    if (k > 0) s += -kx[c1] * g.idx2 * P(c2 - 1);
    s += d * P(c2);
    if (k < g.ncx - 1) s += -kx[c1 + 1] * g.idx2 * P(c2 + 1);
    if (l < g.ncz - 1) s += -kz[c3 + g.nxvz] * g.idz2 * P(c2 + g.ncx);
    return s;
}

struct th_in {
    const double *rho, *Cp, *Qr, *Wd, *Ha, *kx, *kz, *BV;
    const signed char *BT;
};

// b, diag, M^-1, r = b - A x, z = M^-1 r, p = z; partials of r.z and r.r.
__global__ void __launch_bounds__(TB) k_th_setup(th_geom g, double dt, th_in in, const double *__restrict__ x,
                                                 double *__restrict__ diag, double *__restrict__ minv,
                                                 double *__restrict__ r, double *__restrict__ z,
                                                 double *__restrict__ p, double *__restrict__ part_rz,
                                                 double *__restrict__ part_rr) {
    __shared__ double s[TB];
    const int n = g.ncx * g.ncz, blk = xcd_block(blockIdx.x, gridDim.x);
    const int c2 = blk * TB + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (c2 < n) {
        const int k = c2 % g.ncx, l = c2 / g.ncx;
        const double rhoCp = in.rho[c2] * in.Cp[c2];
        double d, bbc;
        th_diag(g, rhoCp / dt, in.kx, in.kz, in.BT, in.BV, k, l, &d, &bbc);
        double b = rhoCp * x[c2] / dt;
        b += 1.0 * (in.Qr ? in.Qr[c2] : 0.0);
        if (in.Wd) b += in.Wd[c2];
        if (in.Ha) b += in.Ha[c2];
        b += 1.0 * bbc;
        const double mi = (fabs(d) > 1e-30) ? 1.0 / d : 1.0;
        const double ri = b - th_row(g, in.kx, in.kz, d, k, l, c2, [&](int j) { return x[j]; });
        diag[c2] = d;
        minv[c2] = mi;
        r[c2] = ri;
        const double zi = mi * ri;
        z[c2] = zi;
        p[c2] = zi;
        rz = ri * zi;
        rr = ri * ri;
    }
    const double a = block_sum(rz, s);
    __syncthreads();
    const double b2 = block_sum(rr, s);
    if (threadIdx.x == 0) {
        part_rz[blk] = a;
        part_rr[blk] = b2;
    }
}

__device__ inline double sum_parts(int nb, const double *__restrict__ part, double *s) {
    double acc = 0.0;
    for (int i = threadIdx.x; i < nb; i += TB) acc += part[i];
    return block_sum(acc, s);
}

__global__ void __launch_bounds__(TB) k_th_fin0(int nb, const double *__restrict__ part_rz,
                                                const double *__restrict__ part_rr, th_ctl *ctl) {
    __shared__ double s[TB];
    const double rz = sum_parts(nb, part_rz, s);
    __syncthreads();
    const double rr = sum_parts(nb, part_rr, s);
    if (threadIdx.x == 0) {
        const double r0 = sqrt(rr);
        ctl->rz = rz;
        ctl->r0 = r0;
        ctl->beta = 0.0;
        ctl->it = 0;
        ctl->done = (r0 < 1e-30) ? 1 : 0;
        ctl->its = (r0 < 1e-30) ? 0 : -1;
    }
}

__global__ void __launch_bounds__(TB) k_th_apply(th_geom g, const double *__restrict__ kx,
                                                 const double *__restrict__ kz, const double *__restrict__ diag,
                                                 const double *__restrict__ z, const double *__restrict__ pold,
                                                 double *__restrict__ pnew, double *__restrict__ q,
                                                 double *__restrict__ part, const th_ctl *__restrict__ ctl) {
    if (ctl->done) return;
    __shared__ double s[TB];
    const double beta = ctl->beta;
    const int n = g.ncx * g.ncz, blk = xcd_block(blockIdx.x, gridDim.x);
    const int c2 = blk * TB + threadIdx.x;
    double pq = 0.0;
    if (c2 < n) {
        const int k = c2 % g.ncx, l = c2 / g.ncx;
        auto P = [&](int j) { return z[j] + beta * pold[j]; };
        const double qi = th_row(g, kx, kz, diag[c2], k, l, c2, P);
        const double pi = P(c2);
        pnew[c2] = pi;
        q[c2] = qi;
        pq = pi * qi;
    }
    const double t = block_sum(pq, s);
    if (threadIdx.x == 0) part[blk] = t;
}

__global__ void __launch_bounds__(TB) k_th_fin1(int nb, const double *__restrict__ part, th_ctl *ctl) {
    if (ctl->done) return;
    __shared__ double s[TB];
    const double pq = sum_parts(nb, part, s);
    if (threadIdx.x == 0) {
        if (fabs(pq) < 1e-30) {
            ctl->its = -2;
            ctl->done = 1;
        } else {
            ctl->alpha = ctl->rz / pq;
        }
    }
}

__global__ void __launch_bounds__(TB) k_th_update(int n, const double *__restrict__ p, const double *__restrict__ q,
                                                  const double *__restrict__ minv, double *__restrict__ x,
                                                  double *__restrict__ r, double *__restrict__ z,
                                                  double *__restrict__ part_rr, double *__restrict__ part_rz,
                                                  const th_ctl *__restrict__ ctl) {
    if (ctl->done) return;
    __shared__ double s[TB];
    const double alpha = ctl->alpha;
    const int blk = xcd_block(blockIdx.x, gridDim.x);
    const int i = blk * TB + threadIdx.x;
    double rr = 0.0, rz = 0.0;
    if (i < n) {
        x[i] += alpha * p[i];
        const double ri = r[i] - alpha * q[i];
        r[i] = ri;
        const double zi = minv[i] * ri;
        z[i] = zi;
        rr = ri * ri;
        rz = ri * zi;
    }
    const double a = block_sum(rr, s);
    __syncthreads();
    const double b = block_sum(rz, s);
    if (threadIdx.x == 0) {
        part_rr[blk] = a;
        part_rz[blk] = b;
    }
}

__global__ void __launch_bounds__(TB) k_th_fin2(int nb, const double *__restrict__ part_rr,
                                                const double *__restrict__ part_rz, th_ctl *ctl) {
    if (ctl->done) return;
    __shared__ double s[TB];
    const double rr = sum_parts(nb, part_rr, s);
    __syncthreads();
    const double rzn = sum_parts(nb, part_rz, s);
    if (threadIdx.x == 0) {
        const int it = ctl->it + 1;
        ctl->it = it;
        if (sqrt(rr) / ctl->r0 < ctl->rel_tol) {
            ctl->its = it;
            ctl->done = 1;
        } else if (fabs(ctl->rz) < 1e-30) {
            ctl->its = -3;
            ctl->done = 1;
        } else {
            ctl->beta = rzn / ctl->rz;
            ctl->rz = rzn;
            if (it >= ctl->max_its) ctl->done = 1;   // its stays -1
        }
    }
}

}  // namespace

struct thermal_scratch {
    dbuf<double> in[8];   // staged inputs: rho, Cp, Qr, Wdiss, Ha, kx, kz, BCT_v
    dbuf<signed char> bt;
    dbuf<double> x, diag, minv, r, z, p[2], q, part[2];
    dbuf<th_ctl> ctl;
};

void mdg_thermal_release(mdg_ctx *c) {
    if (!c->thermal) return;
    thermal_scratch *s = c->thermal;
    for (auto &b : s->in) b.release();
    s->bt.release(); s->x.release(); s->diag.release(); s->minv.release(); s->r.release(); s->z.release();
    s->p[0].release(); s->p[1].release(); s->q.release(); s->part[0].release(); s->part[1].release();
    s->ctl.release();
    delete s;
    c->thermal = nullptr;
}

extern "C" int mdg_energy_step(mdg_ctx *c, double dt, const double *rho_n, const double *Cp, const double *Qr,
                               const double *Wdiss, const double *Ha, const double *kx, const double *kz,
                               const signed char *BCT_t, const double *BCT_v, int max_its, double rel_tol, double *T,
                               int *its) {
    if (!c || !rho_n || !Cp || !kx || !kz || !BCT_t || !BCT_v || !T || !its)
        return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!(dt > 0.0) || max_its < 0 || !(rel_tol > 0.0))
        return mdg_fail(c, MDG_ERR_ARG, "dt and rel_tol must be positive, max_its non-negative");
    if (c->Ncx < 1 || c->Ncz < 1) return mdg_fail(c, MDG_ERR_STATE, "no grid");
    if (!c->thermal) c->thermal = new thermal_scratch();
    thermal_scratch &s = *c->thermal;
    const int ncx = c->Ncx, ncz = c->Ncz, n = ncx * ncz;
    const size_t nkx = (size_t)c->Nx * (c->Nz + 1), nkz = (size_t)(c->Nx + 1) * c->Nz;
    const size_t nbc = (size_t)(ncx + 2) * (ncz + 2);
    th_in in{};
    int rc;
    if ((rc = mdg_stage_in(c, rho_n, (size_t)n, s.in[0], &in.rho)) || (rc = mdg_stage_in(c, Cp, (size_t)n, s.in[1], &in.Cp)) ||
        (rc = mdg_stage_in(c, Qr, (size_t)n, s.in[2], &in.Qr)) ||
        (rc = mdg_stage_in(c, Wdiss, (size_t)n, s.in[3], &in.Wd)) ||
        (rc = mdg_stage_in(c, Ha, (size_t)n, s.in[4], &in.Ha)) || (rc = mdg_stage_in(c, kx, nkx, s.in[5], &in.kx)) ||
        (rc = mdg_stage_in(c, kz, nkz, s.in[6], &in.kz)) || (rc = mdg_stage_in(c, BCT_v, nbc, s.in[7], &in.BV)) ||
        (rc = mdg_stage_in(c, BCT_t, nbc, s.bt, &in.BT)))
        return rc;
    double *x = T;
    if (!mdg_is_device_ptr(T)) {
        HIPCHK(c, s.x.reserve((size_t)n));
        HIPCHK(c, hipMemcpyAsync(s.x.p, T, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
        x = s.x.p;
    }
    const int nb = mdg_grid1d(n, TB);
    HIPCHK(c, s.diag.reserve(n)); HIPCHK(c, s.minv.reserve(n)); HIPCHK(c, s.r.reserve(n));
    HIPCHK(c, s.z.reserve(n)); HIPCHK(c, s.p[0].reserve(n)); HIPCHK(c, s.p[1].reserve(n));
    HIPCHK(c, s.q.reserve(n)); HIPCHK(c, s.part[0].reserve(nb)); HIPCHK(c, s.part[1].reserve(nb));
    HIPCHK(c, s.ctl.reserve(1));
    th_ctl h{};
    h.rel_tol = rel_tol;
    h.max_its = max_its;
    h.done = 0;
    h.its = -1;
    HIPCHK(c, hipMemcpyAsync(s.ctl.p, &h, sizeof h, hipMemcpyHostToDevice, c->stream));
    const th_geom g{ncx, ncz, c->Nx, c->Nx + 1, 1.0 / (c->dx * c->dx), 1.0 / (c->dz * c->dz), c->dx};
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    hipLaunchKernelGGL(k_th_setup, dim3(nb), dim3(TB), 0, c->stream, g, dt, in, (const double *)x, s.diag.p,
                       s.minv.p, s.r.p, s.z.p, s.p[0].p, s.part[0].p, s.part[1].p);
    hipLaunchKernelGGL(k_th_fin0, dim3(1), dim3(TB), 0, c->stream, nb, (const double *)s.part[0].p,
                       (const double *)s.part[1].p, s.ctl.p);
    HIPCHK(c, hipGetLastError());
    // With max_its = 0 the reference's loop does not run: its = -1 unless
    // the start already has a zero residual.
    if (max_its == 0) {
        HIPCHK(c, hipMemcpyAsync(&h, s.ctl.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    int launched = 0, batch = 8, cur = 0;
    while (launched < max_its) {
        const int m = std::min(batch, max_its - launched);
        for (int i = 0; i < m; i++) {
            hipLaunchKernelGGL(k_th_apply, dim3(nb), dim3(TB), 0, c->stream, g, in.kx, in.kz,
                               (const double *)s.diag.p, (const double *)s.z.p, (const double *)s.p[cur].p,
                               s.p[cur ^ 1].p, s.q.p, s.part[0].p, (const th_ctl *)s.ctl.p);
            hipLaunchKernelGGL(k_th_fin1, dim3(1), dim3(TB), 0, c->stream, nb, (const double *)s.part[0].p,
                               s.ctl.p);
            hipLaunchKernelGGL(k_th_update, dim3(nb), dim3(TB), 0, c->stream, n, (const double *)s.p[cur ^ 1].p,
                               (const double *)s.q.p, (const double *)s.minv.p, x, s.r.p, s.z.p, s.part[0].p,
                               s.part[1].p, (const th_ctl *)s.ctl.p);
            hipLaunchKernelGGL(k_th_fin2, dim3(1), dim3(TB), 0, c->stream, nb, (const double *)s.part[0].p,
                               (const double *)s.part[1].p, s.ctl.p);
            cur ^= 1;
        }
        HIPCHK(c, hipGetLastError());
        launched += m;
        HIPCHK(c, hipMemcpyAsync(&h, s.ctl.p, sizeof h, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (h.done) break;
        batch = std::min(batch * 2, 128);
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    if (x != T) HIPCHK(c, hipMemcpyAsync(T, x, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    *its = h.its;
    return MDG_OK;
}
