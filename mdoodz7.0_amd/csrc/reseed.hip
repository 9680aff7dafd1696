// Marker re-seeding of CountPartCell_OLD (ParticleRoutines.c:3224-4050,
// reseed_mode = 0) on the GPU, for a box without free surface and without
// periodic x.
//
// The reference covers the box plus one coarse cell of margin with fine
// cells of dx/2 x dz/2, counts the active markers of each fine cell
// (nearest fine-cell centre), and gives every interior fine cell of a
// non-air coarse cell that holds fewer than floor(min_part_cell / 4)
// markers one new marker at its centre, copying the closest marker among
// the 5 x 5 fine cells around it (first in the reference's list order on
// equal distances).  New markers, taken in x-major fine-cell order, fill the
// inactive slots (phase -1) in increasing order, then are appended.
//
// GPU form: marker counts by atomic increments (exact integers), per-cell
// marker lists by a scan of the counts (order inside a list is arbitrary,
// so the closest marker is the lexicographic minimum of (distance, window
// rank, marker index), which is the reference's first-in-list-order
// choice), the deficient cells and the inactive slots ranked by scans.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "mdg_internal.h"

namespace {

struct rs_geom {
    int Ncx, Ncz, nfx, nfz;
    double dx, dz;
    double fx0, fz0;       // origins of the fine-centre lattice: xc_e[0] - dx/4, zc_e[0] - dz/4
    double xg_e0, zg_e0;   // first extended vertex: xg[0] - dx, zg[0] - dz
    double xg0, zg0, xc0, zc0;
    double span;           // xmax - xmin, FindClosestPhase2's starting distance
};

// (dk, dl) of the 25 fine cells in the order the reference lists their markers
// (ParticleRoutines.c:3783-3808 and the copies at :3820-3951).
__constant__ signed char c_win[25][2] = {
    {0, 0},   {-1, 0},  {1, 0},   {0, -1},  {0, 1},   {-1, -1}, {1, -1},  {-1, 1},  {1, 1},
    {-2, 0},  {2, 0},   {0, -2},  {0, 2},   {-2, -2}, {2, -2},  {-2, 2},  {2, 2},   {-1, -2},
    {1, -2},  {-1, 2},  {1, 2},   {-2, -1}, {-2, 1},  {2, -1},  {2, 1}};

__device__ inline int fine_index(double pos, double origin, double h, int n) {
    const double distance = pos - origin;
    int i = (int)ceil((distance / h * 2.0) + 0.5) - 1;
    i = i < 0 ? 0 : i;
    return i > n - 1 ? n - 1 : i;
}

__global__ void k_rs_count(int n, rs_geom g, const double *__restrict__ x, const double *__restrict__ z,
                           const int *__restrict__ phase, int *__restrict__ cell, int *__restrict__ cnt,
                           int *__restrict__ inactive) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool off = phase[i] == -1;
    inactive[i] = off ? 1 : 0;
    if (off) { cell[i] = -1; return; }
    const int ic = fine_index(x[i], g.fx0, g.dx, g.nfx);
    const int jc = fine_index(z[i], g.fz0, g.dz, g.nfz);
    const int kc = ic + jc * g.nfx;
    cell[i] = kc;
    atomicAdd(&cnt[kc], 1);
}

__global__ void k_rs_fill(int n, const int *__restrict__ cell, const int *__restrict__ off, int *__restrict__ fill,
                          int *__restrict__ lst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int kc = cell[i];
    if (kc < 0) return;
    lst[off[kc] + atomicAdd(&fill[kc], 1)] = i;
}

// Interior fine cell q (x-major: q = (k - 2) (nfz - 4) + (l - 2)) -> (k, l).
__device__ inline void interior_kl(const rs_geom &g, int q, int *k, int *l) {
    const int nz_in = g.nfz - 4;
    *k = q / nz_in + 2;
    *l = q % nz_in + 2;
}

__global__ void k_rs_flag(int nq, rs_geom g, int thr, const signed char *__restrict__ tag, const int *__restrict__ cnt,
                          int *__restrict__ flag) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq) return;
    int k, l;
    interior_kl(g, q, &k, &l);
    double distance = (g.xg_e0 + g.dx / 4.0 + (k)*g.dx / 2.0) - g.xc0;
    int ic = (int)ceil((distance / g.dx) + 0.5) - 1;
    ic = ic < 0 ? 0 : (ic > g.Ncx - 1 ? g.Ncx - 1 : ic);
    distance = (g.zg_e0 + g.dz / 4.0 + (l)*g.dz / 2.0) - g.zc0;
    int jc = (int)ceil((distance / g.dz) + 0.5) - 1;
    jc = jc < 0 ? 0 : (jc > g.Ncz - 1 ? g.Ncz - 1 : jc);
    flag[q] = ((!tag || tag[ic + jc * g.Ncx] != 30) && cnt[k + l * g.nfx] < thr) ? 1 : 0;
}

__global__ void k_rs_new(int nq, rs_geom g, const int *__restrict__ flag, const int *__restrict__ rank,
                         const int *__restrict__ off, const int *__restrict__ lst, const double *__restrict__ x,
                         const double *__restrict__ z, double *__restrict__ nx, double *__restrict__ nz,
                         int *__restrict__ src, int *__restrict__ err) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nq || !flag[q]) return;
    int k, l;
    interior_kl(g, q, &k, &l);
    const int kd = k + l * g.nfx;
    const double new_x = g.xg0 + (k - 2) * g.dx / 2.0 + g.dx / 4.0;
    const double new_z = g.zg0 + (l - 2) * g.dz / 2.0 + g.dz / 4.0;
    double best_d = g.span;
    int best_w = 26, best_i = 0;      // best candidate with distance < span
    int first_w = 26, first_i = 0;    // first marker in list order (the fallback)
    for (int w = 0; w < 25; w++) {
        const int c = kd + c_win[w][0] + c_win[w][1] * g.nfx;
        for (int m = off[c]; m < off[c + 1]; m++) {
            const int i = lst[m];
            if (w < first_w || (w == first_w && i < first_i)) { first_w = w; first_i = i; }
            const double ddx = new_x - x[i], ddz = new_z - z[i];
            const double dst = sqrt(ddx * ddx + ddz * ddz);
            if (dst < best_d || (dst == best_d && best_w < 26 && (w < best_w || (w == best_w && i < best_i)))) {
                best_d = dst;
                best_w = w;
                best_i = i;
            }
        }
    }
    if (first_w == 26) { atomicExch(err, 1); return; }
    const int j = rank[q];
    nx[j] = new_x;
    nz[j] = new_z;
    src[j] = best_w < 26 ? best_i : first_i;
}

__global__ void k_rs_write(int nnew, int n, int npreuse, const int *__restrict__ reuse, const double *__restrict__ nx,
                           const double *__restrict__ nz, const int *__restrict__ src, double *__restrict__ x,
                           double *__restrict__ z, int *__restrict__ phase, int *__restrict__ new_slot,
                           int *__restrict__ new_src) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nnew) return;
    const int s = j < npreuse ? reuse[j] : n + (j - npreuse);
    x[s] = nx[j];
    z[s] = nz[j];
    phase[s] = phase[src[j]];   // sources are active markers, slots are not: no overlap
    new_slot[j] = s;
    new_src[j] = src[j];
}

// reuse[pos[i]] = i for the inactive markers (pos = exclusive scan of the flags)
__global__ void k_rs_compact(int n, const int *__restrict__ inactive, const int *__restrict__ pos,
                             int *__restrict__ reuse) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && inactive[i]) reuse[pos[i]] = i;
}

template <typename T>
__global__ void k_rs_zero(long n, T *__restrict__ a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = T(0);
}

}  // namespace

struct reseed_scratch {
    dbuf<int> cell, cnt, off, fill, lst, inactive, ipos, reuse, flag, rank, src;
    dbuf<double> nx, nz, stage_x, stage_z;
    dbuf<int> stage_p, stage_slot, stage_src, err;
    dbuf<unsigned char> tmp;
    dbuf<signed char> tag;
};

void mdg_reseed_release(mdg_ctx *c) {
    if (!c->reseed) return;
    reseed_scratch *s = c->reseed;
    for (dbuf<int> *b : {&s->cell, &s->cnt, &s->off, &s->fill, &s->lst, &s->inactive, &s->ipos, &s->reuse, &s->flag,
                         &s->rank, &s->src, &s->stage_p, &s->stage_slot, &s->stage_src, &s->err})
        b->release();
    s->nx.release(); s->nz.release(); s->stage_x.release(); s->stage_z.release(); s->tmp.release(); s->tag.release();
    delete s;
    c->reseed = nullptr;
}

// out[i] = sum of in[0..i) for i <= n (out has n + 1 entries)
static int scan_excl(mdg_ctx *c, reseed_scratch &s, const int *in, int *out, int n) {
    size_t bytes = 0;
    HIPCHK(c, rocprim::exclusive_scan(nullptr, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), c->stream));
    HIPCHK(c, s.tmp.reserve(bytes));
    HIPCHK(c, rocprim::exclusive_scan(s.tmp.p, bytes, in, out, 0, (size_t)n, rocprim::plus<int>(), c->stream));
    return MDG_OK;
}

template <typename T>
static void zero(mdg_ctx *c, T *p, long n) {
    if (n > 0) hipLaunchKernelGGL(k_rs_zero<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, p);
}

extern "C" int mdg_count_part_cell(mdg_ctx *c, int min_part_cell, const signed char *cell_tag, int n, int capacity,
                                   double *x, double *z, int *phase, int *new_slot, int *new_src, int *n_new,
                                   int *n_out) {
    if (!c || !x || !z || !phase || !new_slot || !new_src || !n_new || !n_out || n < 0 || capacity < n)
        return mdg_fail(c, MDG_ERR_ARG, "bad argument");
    if (!c->reseed) c->reseed = new reseed_scratch();
    reseed_scratch &s = *c->reseed;
    const signed char *dtag = nullptr;
    int rc;
    if (cell_tag && (rc = mdg_stage_in(c, cell_tag, (size_t)c->Ncx * c->Ncz, s.tag, &dtag))) return rc;
    rs_geom g{};
    g.Ncx = c->Ncx; g.Ncz = c->Ncz;
    g.nfx = 2 * (c->Ncx + 2); g.nfz = 2 * (c->Ncz + 2);
    g.dx = c->dx; g.dz = c->dz;
    g.xg0 = c->h_xg[0]; g.zg0 = c->h_zg[0]; g.xc0 = c->h_xc[0]; g.zc0 = c->h_zc[0];
    g.xg_e0 = g.xg0 - g.dx; g.zg_e0 = g.zg0 - g.dz;
    const double xc_e0 = 0.5 * (g.xg_e0 + (g.xg_e0 + g.dx)), zc_e0 = 0.5 * (g.zg_e0 + (g.zg_e0 + g.dz));
    g.fx0 = xc_e0 - g.dx / 4.0; g.fz0 = zc_e0 - g.dz / 4.0;
    g.span = c->xmax - c->xmin;
    const long nf = (long)g.nfx * g.nfz;
    const int nq = (g.nfx - 4) * (g.nfz - 4);
    const int thr = (int)(ceil((double)min_part_cell) / 4.0);
    // marker arrays: used in place when on the device, staged otherwise
    const bool dev = mdg_is_device_ptr(x) && mdg_is_device_ptr(z) && mdg_is_device_ptr(phase);
    double *dx_ = x, *dz_ = z;
    int *dp_ = phase;
    if (!dev) {
        HIPCHK(c, s.stage_x.reserve((size_t)capacity)); HIPCHK(c, s.stage_z.reserve((size_t)capacity));
        HIPCHK(c, s.stage_p.reserve((size_t)capacity));
        dx_ = s.stage_x.p; dz_ = s.stage_z.p; dp_ = s.stage_p.p;
        if (n > 0) {
            HIPCHK(c, hipMemcpyAsync(dx_, x, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(dz_, z, (size_t)n * 8, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(dp_, phase, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
        }
    }
    const size_t N = (size_t)n + 1;
    HIPCHK(c, s.cell.reserve(N)); HIPCHK(c, s.lst.reserve(N)); HIPCHK(c, s.inactive.reserve(N));
    HIPCHK(c, s.ipos.reserve(N + 1)); HIPCHK(c, s.reuse.reserve(N));
    HIPCHK(c, s.cnt.reserve((size_t)nf + 1)); HIPCHK(c, s.off.reserve((size_t)nf + 2));
    HIPCHK(c, s.fill.reserve((size_t)nf));
    HIPCHK(c, s.flag.reserve((size_t)nq + 1)); HIPCHK(c, s.rank.reserve((size_t)nq + 2));
    HIPCHK(c, s.nx.reserve((size_t)nq)); HIPCHK(c, s.nz.reserve((size_t)nq)); HIPCHK(c, s.src.reserve((size_t)nq));
    HIPCHK(c, s.err.reserve(4));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    zero(c, s.cnt.p, nf + 1);
    zero(c, s.fill.p, nf);
    zero(c, s.err.p, 4);
    zero(c, s.inactive.p, (long)N);
    zero(c, s.flag.p, (long)nq + 1);
    if (n > 0)
        hipLaunchKernelGGL(k_rs_count, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n, g, (const double *)dx_,
                           (const double *)dz_, (const int *)dp_, s.cell.p, s.cnt.p, s.inactive.p);
    if ((rc = scan_excl(c, s, s.cnt.p, s.off.p, (int)nf + 1))) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_rs_fill, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n, (const int *)s.cell.p,
                           (const int *)s.off.p, s.fill.p, s.lst.p);
    hipLaunchKernelGGL(k_rs_flag, dim3(mdg_grid1d(nq, 256)), dim3(256), 0, c->stream, nq, g, thr,
                       dtag, (const int *)s.cnt.p, s.flag.p);
    if ((rc = scan_excl(c, s, s.flag.p, s.rank.p, nq + 1))) return rc;
    if ((rc = scan_excl(c, s, s.inactive.p, s.ipos.p, n + 1))) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_rs_compact, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n,
                           (const int *)s.inactive.p, (const int *)s.ipos.p, s.reuse.p);
    hipLaunchKernelGGL(k_rs_new, dim3(mdg_grid1d(nq, 256)), dim3(256), 0, c->stream, nq, g, (const int *)s.flag.p,
                       (const int *)s.rank.p, (const int *)s.off.p, (const int *)s.lst.p, (const double *)dx_,
                       (const double *)dz_, s.nx.p, s.nz.p, s.src.p, s.err.p);
    HIPCHK(c, hipGetLastError());
    int h[3] = {0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(&h[0], s.rank.p + nq, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&h[1], s.ipos.p + n, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(&h[2], s.err.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int nnew = h[0], npreuse = h[1];
    if (h[2]) return mdg_fail(c, MDG_ERR_NO_MARKERS, "a marker-deficient cell has no marker in its 5 x 5 neighbourhood");
    const long appended = nnew > npreuse ? (long)nnew - npreuse : 0;
    if ((long)n + appended > (long)capacity)
        return mdg_fail(c, MDG_ERR_ARG, "re-seeding needs %ld marker slots, capacity is %d", (long)n + appended, capacity);
    const bool dev_out = mdg_is_device_ptr(new_slot) && mdg_is_device_ptr(new_src);
    int *ds = new_slot, *dsrc = new_src;
    if (!dev_out) {
        HIPCHK(c, s.stage_slot.reserve((size_t)nq)); HIPCHK(c, s.stage_src.reserve((size_t)nq));
        ds = s.stage_slot.p; dsrc = s.stage_src.p;
    }
    if (nnew > 0)
        hipLaunchKernelGGL(k_rs_write, dim3(mdg_grid1d(nnew, 256)), dim3(256), 0, c->stream, nnew, n, npreuse,
                           (const int *)s.reuse.p, (const double *)s.nx.p, (const double *)s.nz.p,
                           (const int *)s.src.p, dx_, dz_, dp_, ds, dsrc);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    const int nout = n + (int)appended;
    if (!dev) {
        // copy back what changed: the whole marker range (reused slots lie anywhere below n)
        if (nout > 0) {
            HIPCHK(c, hipMemcpyAsync(x, dx_, (size_t)nout * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync(z, dz_, (size_t)nout * 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipMemcpyAsync(phase, dp_, (size_t)nout * 4, hipMemcpyDeviceToHost, c->stream));
        }
    }
    if (!dev_out && nnew > 0) {
        HIPCHK(c, hipMemcpyAsync(new_slot, ds, (size_t)nnew * 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(new_src, dsrc, (size_t)nnew * 4, hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    *n_new = nnew;
    *n_out = nout;
    return MDG_OK;
}
