// Marker storage: sort by (cell, quarter cell), index export, field gather.
//
// Sort key.  For a marker at x the reference's nearest-node rule
// (ParticleRoutines.c:2300-2310) gives a centre column ip_c (nodes xc) and a
// vertex column ip_v (nodes xg).  ip_c + ip_v is the quarter-cell column
// (2*cell + 0 for the west half, + 1 for the east half), so the key built
// from the exact reference indices groups every marker with the markers of
// its centre node and of its vertex node.  Keys are 4*cell + quadrant with
// quadrant = xhalf + 2*zhalf; inactive markers (phase -1) sort last.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "mdg_internal.h"
#include "p2g_common.h"

__global__ void k_marker_keys(int n, const double *__restrict__ x, const double *__restrict__ z,
                              const int *__restrict__ phase, double xc0, double xg0, double zc0,
                              double zg0, double dx, double dz, int Ncx, int Ncz, unsigned sentinel,
                              unsigned *__restrict__ keys, int *__restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    idx[i] = i;
    if (phase[i] == -1) { keys[i] = sentinel; return; }
    const int fx = near_index(x[i], xc0, dx, Ncx) + near_index(x[i], xg0, dx, Ncx + 1);
    const int fz = near_index(z[i], zc0, dz, Ncz) + near_index(z[i], zg0, dz, Ncz + 1);
    const unsigned cell = (unsigned)(fx >> 1) + (unsigned)(fz >> 1) * (unsigned)Ncx;
    keys[i] = 4u * cell + (unsigned)((fx & 1) | ((fz & 1) << 1));
}

__global__ void k_gather_markers(int n, const int *__restrict__ perm, const double *__restrict__ x,
                                 const double *__restrict__ z, const int *__restrict__ phase,
                                 double *__restrict__ mx, double *__restrict__ mz,
                                 int *__restrict__ mp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    mx[i] = x[s];
    mz[i] = z[s];
    mp[i] = phase[s];
}

// qoff[q] = first sorted slot whose key is >= q, for q in [0, Q].
__global__ void k_quarter_offsets(int n, const unsigned *__restrict__ keys, unsigned Q,
                                  unsigned *__restrict__ qoff) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    const unsigned lo = (i == 0) ? 0u : keys[i - 1] + 1u;
    const unsigned hi = (i == n) ? Q : min(keys[i], Q);
    for (unsigned q = lo; q <= hi; q++) qoff[q] = (unsigned)i;
}

__global__ void k_gather_field(int n, const int *__restrict__ perm, const double *__restrict__ src,
                               double *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

__global__ void k_marker_index(int n, const int *__restrict__ perm, const double *__restrict__ mx,
                               const double *__restrict__ mz, double X0, double Z0, double dx,
                               double dz, int nx, int nz, int *__restrict__ ip, int *__restrict__ jp) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = perm[i];
    // the same index routine as the P2G kernels, so the bit-exact index
    // checks (ties, edges) cover the division-free path
    ip[s] = near_index_fast(mx[i], X0, dx, 1.0 / dx, nx);
    jp[s] = near_index_fast(mz[i], Z0, dz, 1.0 / dz, nz);
}

extern "C" int mdg_markers_set(mdg_ctx *c, int n, const double *x, const double *z,
                               const int *phase) {
    if (!c || n < 0 || (n > 0 && (!x || !z || !phase))) return mdg_fail(c, MDG_ERR_ARG, "bad markers");
    const size_t N = (size_t)n;
    const double *dx_, *dz_;
    const int *dp_;
    int rc;
    // The previous marker set is invalid from here on: a failure below may
    // already have reallocated the sorted buffers, so no cached weights or
    // offsets may be used until this call succeeds.
    c->n_markers = 0;
    for (int q = 0; q < 4; q++) c->mw_state[q] = 0;
    HIPCHK(c, c->field_tmp.reserve(2 * N + 1));
    HIPCHK(c, c->idx.reserve(N + 1));
    scoped_dbuf<double> tx, tz;  // only used when host data must be staged
    scoped_dbuf<int> tp;
    if ((rc = mdg_stage_in(c, x, N, tx, &dx_)) || (rc = mdg_stage_in(c, z, N, tz, &dz_)) ||
        (rc = mdg_stage_in(c, phase, N, tp, &dp_)))
        return rc;
    const unsigned ncells = (unsigned)c->Ncx * (unsigned)c->Ncz;
    const unsigned Q = 4u * ncells;
    HIPCHK(c, c->keys.reserve(N + 1));
    HIPCHK(c, c->keys2.reserve(N + 1));
    HIPCHK(c, c->perm.reserve(N + 1));
    HIPCHK(c, c->mx.reserve(N + 1));
    HIPCHK(c, c->mz.reserve(N + 1));
    HIPCHK(c, c->mphase.reserve(N + 1));
    HIPCHK(c, c->qoff.reserve((size_t)Q + 1));
    if (n > 0) {
        hipLaunchKernelGGL(k_marker_keys, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n,
                           dx_, dz_, dp_, c->h_xc[0], c->h_xg[0], c->h_zc[0], c->h_zg[0], c->dx,
                           c->dz, c->Ncx, c->Ncz, Q, c->keys.p, c->idx.p);
        unsigned end_bit = 1;
        while ((1ull << end_bit) <= (unsigned long long)Q) end_bit++;
        size_t tmp_bytes = 0;
        HIPCHK(c, rocprim::radix_sort_pairs(nullptr, tmp_bytes, c->keys.p, c->keys2.p, c->idx.p,
                                            c->perm.p, (unsigned)n, 0u, end_bit, c->stream));
        HIPCHK(c, c->sort_tmp.reserve(tmp_bytes));
        HIPCHK(c, rocprim::radix_sort_pairs(c->sort_tmp.p, tmp_bytes, c->keys.p, c->keys2.p,
                                            c->idx.p, c->perm.p, (unsigned)n, 0u, end_bit,
                                            c->stream));
        hipLaunchKernelGGL(k_gather_markers, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n,
                           c->perm.p, dx_, dz_, dp_, c->mx.p, c->mz.p, c->mphase.p);
    }
    hipLaunchKernelGGL(k_quarter_offsets, dim3(mdg_grid1d((long)n + 1, 256)), dim3(256), 0,
                       c->stream, n, c->keys2.p, Q, c->qoff.p);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->n_markers = n;   // cached P2G weights stay invalid (state 0) until recomputed
    return MDG_OK;
}

extern "C" int mdg_marker_index(mdg_ctx *c, int centroid, int *ip, int *jp) {
    if (!c || !ip || !jp) return mdg_fail(c, MDG_ERR_ARG, "null output");
    node_set ns;
    int rc = mdg_node_set(c, centroid, &ns);
    if (rc) return rc;
    const int n = c->n_markers;
    if (n == 0) return MDG_OK;
    const bool dev = mdg_is_device_ptr(ip) && mdg_is_device_ptr(jp);
    int *dip = ip, *djp = jp;
    scoped_dbuf<int> t;
    if (!dev) {
        HIPCHK(c, t.reserve(2 * (size_t)n));
        dip = t.p;
        djp = t.p + n;
    }
    hipLaunchKernelGGL(k_marker_index, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n,
                       c->perm.p, c->mx.p, c->mz.p, ns.X0, ns.Z0, c->dx, c->dz, ns.nx, ns.nz, dip,
                       djp);
    HIPCHK(c, hipGetLastError());
    if (!dev) {
        HIPCHK(c, hipMemcpyAsync(ip, dip, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(jp, djp, n * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}

// Gather a caller-order per-marker array into sorted order on the device.
int mdg_gather_sorted(mdg_ctx *c, const double *src, double *dst) {
    const int n = c->n_markers;
    if (n == 0) return MDG_OK;
    const double *s;
    scoped_dbuf<double> t;
    int rc = mdg_stage_in(c, src, (size_t)n, t, &s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_field, dim3(mdg_grid1d(n, 256)), dim3(256), 0, c->stream, n,
                       c->perm.p, s, dst);
    HIPCHK(c, hipGetLastError());
    if (t.p) HIPCHK(c, hipStreamSynchronize(c->stream));   // before t is freed
    return MDG_OK;
}

extern "C" int mdg_markers_gather(mdg_ctx *c, const double *src, double *dst_sorted) {
    if (!c || !src || !dst_sorted) return mdg_fail(c, MDG_ERR_ARG, "null array");
    if (!mdg_is_device_ptr(dst_sorted)) return mdg_fail(c, MDG_ERR_ARG, "dst must be device memory");
    int rc = mdg_gather_sorted(c, src, dst_sorted);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return MDG_OK;
}
